// Mask passes of grid resampling for gfx950 (SparseGrid.resample,
// svox2/svox2.py:1223-1424): threshold a value volume into a keep mask,
// dilate the mask (dilate_kernel, svox2/csrc/misc_kernel.cu:20-48), and turn
// it into links plus the list of kept cells.
//
// Dilation.  One pass sets a cell when any cell of its 3x3x3 neighbourhood,
// clipped to the grid, is set.  A thread takes 4 consecutive z cells (one
// 32-bit word of mask bytes) when rows are a multiple of 4 long: it ORs the 9
// words of the x/y neighbourhood, then folds in each word's z neighbours by
// byte shifts of the word with the bytes next to it.
//
// Compaction.  Cells are numbered in raster order (or Morton order for a cubic
// power-of-two grid): the kept cell with the k-th smallest order index gets
// row k.  Pass 1 counts the kept cells of each tile of 4096 order indices
// (one wave per 256-cell slice, ballots and popcounts); one workgroup scans the
// tile counts; pass 2 recounts each tile's cells in order, offsets them by the
// tile's prefix and writes the links and the kept-cell list.
#include "common.hpp"
#include "fastdiv.hpp"

#include <climits>

namespace svox2amd {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 4096;  // order indices per compaction tile
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(kBlock) void k_threshold(const float* __restrict__ v, int64_t n,
                                                      float thresh, int s2, int keep_last_z,
                                                      uint8_t* __restrict__ mask) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const bool last_z = keep_last_z && (i % s2) == s2 - 1;
    mask[i] = (v[i] >= thresh || last_z) ? 1 : 0;
}

// One dilation pass, 4 z cells per thread (s2 % 4 == 0, 4-byte aligned mask).
__global__ __launch_bounds__(kBlock) void k_dilate4(const uint32_t* __restrict__ in, int s0,
                                                    int s1, int s2, FastDiv div_w,
                                                    FastDiv div_y, uint32_t* __restrict__ out) {
    const int w2 = s2 / 4;  // words per row
    const int64_t nw = static_cast<int64_t>(s0) * s1 * w2;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= nw) return;
    uint32_t wz, y;
    const uint32_t row = div_w.divmod(static_cast<uint32_t>(i), wz);
    const uint32_t x = div_y.divmod(row, y);
    uint32_t acc = 0, lo = 0, hi = 0;  // OR of the 9 words and of their z neighbours
    for (int dx = -1; dx <= 1; ++dx) {
        const int xx = static_cast<int>(x) + dx;
        if (xx < 0 || xx >= s0) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = static_cast<int>(y) + dy;
            if (yy < 0 || yy >= s1) continue;
            const int64_t r = (static_cast<int64_t>(xx) * s1 + yy) * w2;
            acc |= in[r + wz];
            if (wz > 0) lo |= in[r + wz - 1];
            if (static_cast<int>(wz) + 1 < w2) hi |= in[r + wz + 1];
        }
    }
    // Byte b of the word is cell 4 wz + b (little endian).  Its z-1 neighbour is
    // byte b-1 (byte 3 of the previous word for b = 0), its z+1 neighbour byte
    // b+1 (byte 0 of the next word for b = 3).
    const uint32_t from_lower = (acc << 8) | (lo >> 24);
    const uint32_t from_upper = (acc >> 8) | (hi << 24);
    const uint32_t any = acc | from_lower | from_upper;
    // Bytes are 0 or 1; make each nonzero byte exactly 1.
    out[i] = (any | (any >> 1) | (any >> 2) | (any >> 3) | (any >> 4) | (any >> 5) |
              (any >> 6) | (any >> 7)) & 0x01010101u;
}

// One dilation pass, a cell per thread (any shape).
__global__ __launch_bounds__(kBlock) void k_dilate1(const uint8_t* __restrict__ in, int s0,
                                                    int s1, int s2, FastDiv div_z, FastDiv div_y,
                                                    uint8_t* __restrict__ out) {
    const int64_t n = static_cast<int64_t>(s0) * s1 * s2;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t z, y;
    const uint32_t row = div_z.divmod(static_cast<uint32_t>(i), z);
    const uint32_t x = div_y.divmod(row, y);
    uint8_t v = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        const int xx = static_cast<int>(x) + dx;
        if (xx < 0 || xx >= s0) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = static_cast<int>(y) + dy;
            if (yy < 0 || yy >= s1) continue;
            for (int dz = -1; dz <= 1; ++dz) {
                const int zz = static_cast<int>(z) + dz;
                if (zz < 0 || zz >= s2) continue;
                v |= in[(static_cast<int64_t>(xx) * s1 + yy) * s2 + zz];
            }
        }
    }
    out[i] = v ? 1 : 0;
}

// Every third bit of v gathered into the low bits (inverse of the Morton
// spread of a 10-bit coordinate).
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
    v &= 0x09249249u;
    v = (v ^ (v >> 2)) & 0x030C30C3u;
    v = (v ^ (v >> 4)) & 0x0300F00Fu;
    v = (v ^ (v >> 8)) & 0x030000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}

struct Order {
    int s1, s2;
    int morton;  // 1: order index = Morton code (x bits at 3k+2, y at 3k+1, z at 3k)
};

__device__ __forceinline__ int64_t cell_of(const Order& o, int64_t k) {
    if (!o.morton) return k;
    const uint32_t c = static_cast<uint32_t>(k);
    const uint32_t x = compact3(c >> 2), y = compact3(c >> 1), z = compact3(c);
    return (static_cast<int64_t>(x) * o.s1 + y) * o.s2 + z;
}

// Pass 1: kept cells per tile; each wave counts 64 order indices at a time.
__global__ __launch_bounds__(kBlock) void k_tile_count(const uint8_t* __restrict__ mask,
                                                       int64_t n, Order o,
                                                       uint32_t* __restrict__ counts) {
    __shared__ uint32_t part[kBlock / 64];
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kTile;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t c = 0;
    for (int off = wave * 64; off < kTile; off += kBlock) {
        const int64_t k = t0 + off + lane;
        const bool keep = k < n && mask[cell_of(o, k)] != 0;
        c += static_cast<uint32_t>(__popcll(__ballot(keep)));
    }
    if (lane == 0) part[wave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kBlock / 64; ++w) s += part[w];
        counts[blockIdx.x] = s;
    }
}

// Exclusive scan of the tile counts in one workgroup; counts[n_tiles] gets the total.
__global__ __launch_bounds__(kScanBlock) void k_tile_scan(uint32_t* __restrict__ counts,
                                                          int n_tiles) {
    __shared__ uint32_t sums[kScanBlock];
    const int per = (n_tiles + kScanBlock - 1) / kScanBlock;
    const int lo = threadIdx.x * per, hi = min(lo + per, n_tiles);
    uint32_t s = 0;
    for (int i = lo; i < hi; ++i) s += counts[i];
    sums[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < kScanBlock; d <<= 1) {
        const uint32_t v = threadIdx.x >= static_cast<unsigned>(d) ? sums[threadIdx.x - d] : 0u;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = sums[threadIdx.x] - s;  // exclusive prefix of this thread's range
    for (int i = lo; i < hi; ++i) {
        const uint32_t c = counts[i];
        counts[i] = run;
        run += c;
    }
    if (threadIdx.x == kScanBlock - 1) counts[n_tiles] = sums[kScanBlock - 1];
}

// Pass 2: links and the kept-cell list, tile by tile in order.
__global__ __launch_bounds__(kBlock) void k_tile_assign(const uint8_t* __restrict__ mask,
                                                        int64_t n, Order o,
                                                        const uint32_t* __restrict__ prefix,
                                                        int32_t* __restrict__ links,
                                                        int32_t* __restrict__ kept) {
    __shared__ uint32_t wave_base[kBlock / 64];
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kTile;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t base = prefix[blockIdx.x];
    for (int off = 0; off < kTile; off += kBlock) {
        const int64_t k = t0 + off + threadIdx.x;
        int64_t cell = -1;
        bool keep = false;
        if (k < n) {
            cell = cell_of(o, k);
            keep = mask[cell] != 0;
        }
        const uint64_t b = __ballot(keep);
        if (lane == 0) wave_base[wave] = static_cast<uint32_t>(__popcll(b));
        __syncthreads();
        uint32_t before = base;
        for (int w = 0; w < wave; ++w) before += wave_base[w];
        uint32_t total = 0;
        for (int w = 0; w < kBlock / 64; ++w) total += wave_base[w];
        const uint32_t rank =
            before + static_cast<uint32_t>(__popcll(b & ((uint64_t{1} << lane) - 1)));
        if (cell >= 0) {
            links[cell] = keep ? static_cast<int32_t>(rank) : -1;
            if (keep && kept != nullptr) kept[rank] = static_cast<int32_t>(cell);
        }
        base += total;
        __syncthreads();
    }
}

int check_volume(const int32_t size[3], int64_t& n) {
    SVOX2_REQUIRE(size != nullptr, "size is NULL");
    n = 1;
    for (int i = 0; i < 3; ++i) {
        SVOX2_REQUIRE(size[i] >= 1, "size[%d] = %d", i, size[i]);
        n *= size[i];
    }
    SVOX2_REQUIRE(n < INT32_MAX, "volume has too many cells");
    return ok();
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int make_order(const int32_t size[3], int32_t z_order, Order& o) {
    o.s1 = size[1];
    o.s2 = size[2];
    o.morton = z_order ? 1 : 0;
    if (z_order)
        SVOX2_REQUIRE(size[0] == size[1] && size[1] == size[2] && is_pow2(size[0]) &&
                          size[0] <= 1024,
                      "Morton order needs a cubic power-of-two grid of side <= 1024");
    return ok();
}

int64_t n_tiles_of(int64_t n) { return (n + kTile - 1) / kTile; }

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

extern "C" int svox2_threshold_mask(const float* values, const int32_t size[3], float thresh,
                                    int32_t keep_last_z, uint8_t* mask_out, void* stream) {
    int64_t n;
    int st = check_volume(size, n);
    if (st) return st;
    SVOX2_REQUIRE(values != nullptr && mask_out != nullptr, "values / mask_out are NULL");
    k_threshold<<<blocks_for(n, kBlock), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
        values, n, thresh, size[2], keep_last_z, mask_out);
    return check_launch("threshold_mask");
}

extern "C" int svox2_dilate(const uint8_t* mask, const int32_t size[3], uint8_t* out,
                            void* stream) {
    int64_t n;
    int st = check_volume(size, n);
    if (st) return st;
    SVOX2_REQUIRE(mask != nullptr && out != nullptr, "mask / out are NULL");
    SVOX2_REQUIRE(mask != out, "dilate cannot run in place");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool vec = size[2] % 4 == 0 && reinterpret_cast<uintptr_t>(mask) % 4 == 0 &&
                     reinterpret_cast<uintptr_t>(out) % 4 == 0;
    if (vec) {
        const int64_t nw = n / 4;
        k_dilate4<<<blocks_for(nw, kBlock), kBlock, 0, s>>>(
            reinterpret_cast<const uint32_t*>(mask), size[0], size[1], size[2],
            FastDiv(static_cast<uint32_t>(size[2] / 4)), FastDiv(static_cast<uint32_t>(size[1])),
            reinterpret_cast<uint32_t*>(out));
    } else {
        k_dilate1<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
            mask, size[0], size[1], size[2], FastDiv(static_cast<uint32_t>(size[2])),
            FastDiv(static_cast<uint32_t>(size[1])), out);
    }
    return check_launch("dilate");
}

extern "C" size_t svox2_mask_compact_workspace_bytes(const int32_t size[3]) {
    int64_t n = 1;
    for (int i = 0; i < 3; ++i) n *= size[i] > 0 ? size[i] : 0;
    return static_cast<size_t>(n_tiles_of(n) + 1) * sizeof(uint32_t);
}

extern "C" int svox2_mask_compact_count(const uint8_t* mask, const int32_t size[3],
                                        int32_t z_order, void* workspace,
                                        size_t workspace_bytes, void* stream) {
    int64_t n;
    int st = check_volume(size, n);
    if (st) return st;
    Order o;
    if ((st = make_order(size, z_order, o))) return st;
    SVOX2_REQUIRE(mask != nullptr && workspace != nullptr, "mask / workspace are NULL");
    SVOX2_REQUIRE(workspace_bytes >= svox2_mask_compact_workspace_bytes(size),
                  "workspace holds %zu bytes, needs %zu", workspace_bytes,
                  svox2_mask_compact_workspace_bytes(size));
    const int64_t tiles = n_tiles_of(n);
    uint32_t* counts = static_cast<uint32_t*>(workspace);
    hipStream_t s = static_cast<hipStream_t>(stream);
    k_tile_count<<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(mask, n, o, counts);
    k_tile_scan<<<1, kScanBlock, 0, s>>>(counts, static_cast<int>(tiles));
    return check_launch("mask_compact_count");
}

extern "C" int svox2_mask_compact_assign(const uint8_t* mask, const int32_t size[3],
                                         int32_t z_order, const void* workspace,
                                         int32_t* links_out, int32_t* kept_cells_out,
                                         void* stream) {
    int64_t n;
    int st = check_volume(size, n);
    if (st) return st;
    Order o;
    if ((st = make_order(size, z_order, o))) return st;
    SVOX2_REQUIRE(mask != nullptr && workspace != nullptr && links_out != nullptr,
                  "mask / workspace / links_out are NULL");
    const int64_t tiles = n_tiles_of(n);
    k_tile_assign<<<static_cast<unsigned>(tiles), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
        mask, n, o, static_cast<const uint32_t*>(workspace), links_out, kept_cells_out);
    return check_launch("mask_compact_assign");
}
