// Sparse total-variation gradient for gfx950.
//
// What is computed (tv_grad_sparse_kernel, svox2/csrc/loss_kernel.cu:178-246,
// and its launcher :558-604), for each listed cell (x, y, z) and channel c:
//   v000 = the cell's value, v001 / v010 / v100 = its +z / +y / +x neighbours;
//   dx, dy, dz = neighbour - v000;
//   idelta = (scale / n_cells) / sqrt(1e-9 + dx^2 + dy^2 + dz^2)   (unscaled diffs);
//   dx, dy, dz *= size_axis / 256;
//   grad[v000] += -(dx + dy + dz) idelta, grad[v001] += dz idelta, ... for every
//   link that addresses a row and value that is nonzero, and the row's mask.
// The reference's edge rules are kept: a neighbour past the grid's end reads
// link 0 (row 0) and receives gradient; a missing neighbour reads 0, or v000
// with ignore_edge, which also skips cells whose link is 0; ignore_last_z
// skips cells with z = size2 - 2.
//
// Layout of the work.  A wavefront reads 64 listed cells at a time, one lane
// per cell: it splits the flat index with invariant-divisor multiplies (no
// integer division), reads the cell's four links and decides whether any of
// them addresses a row.  At about 8 % occupancy most cells address none and
// stop there.  With several channels a ballot lists the live cells in LDS,
// and the workgroup's half-waves process them, one 32-lane half per cell and
// lane = channel: the half's four data reads are each one coalesced run of a
// row, and each of its four gradient updates is one atomic instruction into
// one row, so the interleaved 128-byte voxel rows are touched once per cell
// and neighbour.  One-channel data (density) skips the list: each lane
// finishes its own cell.
#include "common.hpp"
#include "fastdiv.hpp"

#include <algorithm>

namespace svox2amd {
namespace {

constexpr int kBlock = 256;

template <typename DT>
struct TvJob {
    const int32_t* links;
    int s0, s1, s2;
    FastDiv div_z;  // by s2
    FastDiv div_y;  // by s1
    const DT* data;
    int rows;
    int64_t dst, gst;  // row strides of data and grad
    const int32_t* cells;  // listed cells, or NULL for the run first_cell + k (mod n_total)
    int64_t n_cells;
    int64_t first_cell;
    int64_t n_total;
    int start_dim, nch;
    float scale;  // already divided by n_cells
    int ignore_edge, ignore_last_z;
    uint8_t* mask;
    float* grad;
    float sc0, sc1, sc2;  // size_axis / 256
};

struct CellLinks {
    int l000, l001, l010, l100;
    bool live;
};

template <typename DT>
__device__ __forceinline__ bool valid_row(const TvJob<DT>& J, int l) {
    return static_cast<uint32_t>(l) < static_cast<uint32_t>(J.rows);
}

// The four links of listed cell k (live = false if it contributes nothing).
template <typename DT>
__device__ __forceinline__ CellLinks cell_links(const TvJob<DT>& J, int64_t k) {
    CellLinks c{-1, -1, -1, -1, false};
    if (k >= J.n_cells) return c;
    int64_t run = J.first_cell + k;
    if (run >= J.n_total) run -= J.n_total;
    const int xyz = J.cells != nullptr ? J.cells[k] : static_cast<int>(run);
    if (xyz < 0) return c;
    uint32_t z, y;
    const uint32_t xy = J.div_z.divmod(static_cast<uint32_t>(xyz), z);
    const uint32_t x = J.div_y.divmod(xy, y);
    if (x >= static_cast<uint32_t>(J.s0)) return c;  // past the grid
    const int64_t base = (static_cast<int64_t>(x) * J.s1 + y) * J.s2 + z;
    const int32_t* lp = J.links + base;
    c.l000 = lp[0];
    if (J.ignore_edge && c.l000 == 0) return c;
    if (J.ignore_last_z && static_cast<int>(z) == J.s2 - 2) return c;
    c.l001 = static_cast<int>(z) + 1 < J.s2 ? lp[1] : 0;
    c.l010 = static_cast<int>(y) + 1 < J.s1 ? lp[J.s2] : 0;
    c.l100 = static_cast<int>(x) + 1 < J.s0 ? lp[static_cast<int64_t>(J.s1) * J.s2] : 0;
    c.live = valid_row(J, c.l000) || valid_row(J, c.l001) || valid_row(J, c.l010) ||
             valid_row(J, c.l100);
    return c;
}

// The four gradient terms of channel ch of a cell, in the reference's float
// operation order, and which of them the reference applies: a term is added
// (and its row marked) when its scaled difference is nonzero, whatever the
// product with idelta rounds to.
template <typename DT>
__device__ __forceinline__ void cell_terms(const TvJob<DT>& J, const CellLinks& c, int ch,
                                           float t[4], bool act[4]) {
    const float v000 = valid_row(J, c.l000) ? ld(J.data, c.l000 * J.dst + ch) : 0.f;
    const float nul = J.ignore_edge ? v000 : 0.f;
    const float v001 = valid_row(J, c.l001) ? ld(J.data, c.l001 * J.dst + ch) : nul;
    const float v010 = valid_row(J, c.l010) ? ld(J.data, c.l010 * J.dst + ch) : nul;
    const float v100 = valid_row(J, c.l100) ? ld(J.data, c.l100 * J.dst + ch) : nul;
    float dx = v100 - v000;
    float dy = v010 - v000;
    float dz = v001 - v000;
    const float idelta = J.scale * rsqrtf(1e-9f + dx * dx + dy * dy + dz * dz);
    dx *= J.sc0;
    dy *= J.sc1;
    dz *= J.sc2;
    const float sm = -(dx + dy + dz);
    const float v[4] = {sm, dz, dy, dx};
    const int ls[4] = {c.l000, c.l001, c.l010, c.l100};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        t[i] = v[i] * idelta;
        act[i] = valid_row(J, ls[i]) && v[i] != 0.f;
    }
}

// One channel (density): a wavefront takes 64 listed cells at a time and
// each lane finishes its own cell.
template <typename DT>
__global__ __launch_bounds__(kBlock) void k_tv_cells(TvJob<DT> J) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * kBlock) >> 6;
    for (int64_t base = wave * 64; base < J.n_cells; base += nwaves * 64) {
        const CellLinks c = cell_links(J, base + lane);
        if (!c.live) continue;
        float t[4];
        bool act[4];
        const int ch = J.start_dim;
        cell_terms(J, c, ch, t, act);
        const int ls[4] = {c.l000, c.l001, c.l010, c.l100};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (act[i]) {
                atomic_add_f32(J.grad + ls[i] * J.gst + ch, t[i]);
                if (J.mask) J.mask[ls[i]] = 1;
            }
        }
    }
}

// Several channels (SH): a workgroup takes kTvWindow listed cells.  Its first
// wavefront reads their links and lists the live ones in LDS (ballot
// compaction); then the workgroup's 8 half-waves take the listed cells in
// turn, lane = channel.  Live cells come in runs (a contiguous run of cells
// crossing the occupied region), and spreading a run over 8 half-waves
// shortens the longest chain of dependent row reads and atomics.
constexpr int kTvWindow = 64;

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_tv_cells_multi(TvJob<DT> J) {
    __shared__ int4 live[kTvWindow];
    __shared__ int n_live;
    const int hw = threadIdx.x >> 5, hl = threadIdx.x & 31;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kTvWindow; base < J.n_cells;
         base += static_cast<int64_t>(gridDim.x) * kTvWindow) {
        if (threadIdx.x < 64) {
            const CellLinks c = cell_links(J, base + threadIdx.x);
            const uint64_t m = __ballot(c.live);
            if (c.live)
                live[__popcll(m & ((uint64_t{1} << threadIdx.x) - 1))] =
                    make_int4(c.l000, c.l001, c.l010, c.l100);
            if (threadIdx.x == 0) n_live = static_cast<int>(__popcll(m));
        }
        __syncthreads();
        const int total = n_live;
        for (int j = hw; j < total; j += kBlock / 32) {
            const int4 L = live[j];
            const CellLinks h{L.x, L.y, L.z, L.w, true};
            const int ls[4] = {L.x, L.y, L.z, L.w};
            bool any[4] = {false, false, false, false};
            for (int k = hl; k < J.nch; k += 32) {
                const int ch = J.start_dim + k;
                float t[4];
                bool act[4];
                cell_terms(J, h, ch, t, act);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (act[i]) {
                        atomic_add_f32(J.grad + ls[i] * J.gst + ch, t[i]);
                        any[i] = true;
                    }
                }
            }
            if (J.mask) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    // one store per half-wave and row that received a gradient
                    const uint64_t b = __ballot(any[i]);
                    const uint32_t mine = static_cast<uint32_t>(b >> (32 * (hw & 1)));
                    if (mine != 0 && hl == __builtin_ctz(mine)) J.mask[ls[i]] = 1;
                }
            }
        }
        __syncthreads();  // the list is rewritten by the next window
    }
}

template <typename DT>
int launch(const TvJob<DT>& J, hipStream_t s) {
    if (J.nch == 1) {
        const int64_t windows = (J.n_cells + 63) / 64;
        const int64_t blocks = std::min<int64_t>((windows + 3) / 4, 256 * 8);
        k_tv_cells<DT><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(J);
    } else {
        const int64_t windows = (J.n_cells + kTvWindow - 1) / kTvWindow;
        const int64_t blocks = std::min<int64_t>(windows, 256 * 32);
        k_tv_cells_multi<DT><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(J);
    }
    return check_launch("tv_grad_sparse");
}

template <typename DT>
int run_tv(const DT* data, const int32_t* links, const int32_t size[3], int64_t rows,
           int32_t data_dim, int64_t data_stride, int64_t grad_stride, const int32_t* cells,
           int64_t n_cells, int64_t first_cell, uint8_t* mask, int32_t start_dim,
           int32_t end_dim, float scale, int32_t ignore_edge, int32_t ignore_last_z, float* grad,
           hipStream_t s) {
    TvJob<DT> J;
    J.links = links;
    J.s0 = size[0];
    J.s1 = size[1];
    J.s2 = size[2];
    J.div_z = FastDiv(static_cast<uint32_t>(size[2]));
    J.div_y = FastDiv(static_cast<uint32_t>(size[1]));
    J.data = data;
    J.rows = static_cast<int>(rows);
    J.dst = data_stride > 0 ? data_stride : data_dim;
    J.gst = grad_stride > 0 ? grad_stride : data_dim;
    SVOX2_REQUIRE(J.dst >= data_dim && J.gst >= data_dim, "row stride below data_dim");
    J.cells = cells;
    J.n_cells = n_cells;
    J.first_cell = first_cell;
    J.n_total = static_cast<int64_t>(size[0]) * size[1] * size[2];
    J.start_dim = start_dim;
    J.nch = end_dim - start_dim;
    // The reference divides the scale by the number of cells (loss_kernel.cu:597).
    J.scale = scale / static_cast<float>(n_cells);
    J.ignore_edge = ignore_edge;
    J.ignore_last_z = ignore_last_z;
    J.mask = mask;
    J.grad = grad;
    J.sc0 = static_cast<float>(size[0]) * (1.f / 256.f);
    J.sc1 = static_cast<float>(size[1]) * (1.f / 256.f);
    J.sc2 = static_cast<float>(size[2]) * (1.f / 256.f);
    return launch(J, s);
}

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

namespace svox2amd {
namespace {

// Checks shared by both entry points; dispatches on the storage kind.
int tv_entry(const int32_t* links, const int32_t size[3], const void* data, int32_t storage,
             int64_t rows, int32_t data_dim, int64_t data_stride, int64_t grad_stride,
             const int32_t* cells, int64_t n_cells, int64_t first_cell, uint8_t* mask_out,
             int32_t start_dim, int32_t end_dim, float scale, int32_t ignore_edge,
             int32_t ignore_last_z, float* grad_data, void* stream) {
    SVOX2_REQUIRE(size != nullptr && links != nullptr, "links / size are NULL");
    for (int i = 0; i < 3; ++i) SVOX2_REQUIRE(size[i] >= 1, "grid size[%d] = %d", i, size[i]);
    const int64_t cells_total = static_cast<int64_t>(size[0]) * size[1] * size[2];
    SVOX2_REQUIRE(cells_total < INT32_MAX, "too many cells");
    SVOX2_REQUIRE(storage == SVOX2_STORAGE_F32 || storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", storage);
    SVOX2_REQUIRE(rows >= 0 && rows < INT32_MAX, "rows out of range");
    SVOX2_REQUIRE(data_dim >= 1, "data_dim must be >= 1");
    SVOX2_REQUIRE(0 <= start_dim && start_dim < end_dim && end_dim <= data_dim,
                  "channel range [%d, %d) invalid for %d channels", start_dim, end_dim, data_dim);
    SVOX2_REQUIRE(n_cells >= 0, "negative cell count");
    if (cells == nullptr)
        SVOX2_REQUIRE(0 <= first_cell && first_cell < cells_total && n_cells <= cells_total,
                      "cell run [%lld, +%lld) invalid for %lld cells",
                      static_cast<long long>(first_cell), static_cast<long long>(n_cells),
                      static_cast<long long>(cells_total));
    if (n_cells == 0 || rows == 0) return ok();
    SVOX2_REQUIRE(data != nullptr && grad_data != nullptr, "data / grad_data are NULL");
    SVOX2_REQUIRE(data_stride >= 0 && grad_stride >= 0, "negative row stride");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (storage == SVOX2_STORAGE_BF16)
        return run_tv(static_cast<const uint16_t*>(data), links, size, rows, data_dim, data_stride,
                      grad_stride, cells, n_cells, first_cell, mask_out, start_dim, end_dim,
                      scale, ignore_edge, ignore_last_z, grad_data, s);
    return run_tv(static_cast<const float*>(data), links, size, rows, data_dim, data_stride,
                  grad_stride, cells, n_cells, first_cell, mask_out, start_dim, end_dim, scale,
                  ignore_edge, ignore_last_z, grad_data, s);
}

}  // namespace
}  // namespace svox2amd

extern "C" int svox2_tv_grad_sparse(const int32_t* links, const int32_t size[3], const void* data,
                                    int32_t storage, int64_t rows, int32_t data_dim,
                                    int64_t data_stride, int64_t grad_stride,
                                    const int32_t* rand_cells, int64_t n_cells,
                                    uint8_t* mask_out, int32_t start_dim, int32_t end_dim,
                                    float scale, int32_t use_logalpha, float logalpha_delta,
                                    int32_t ignore_edge, int32_t ignore_last_z,
                                    float ndc_coeffx, float ndc_coeffy, float* grad_data,
                                    void* stream) {
    (void)logalpha_delta;
    (void)ndc_coeffx;
    (void)ndc_coeffy;
    SVOX2_REQUIRE(!use_logalpha, "use_logalpha is no longer supported (as in the reference)");
    SVOX2_REQUIRE(n_cells == 0 || rand_cells != nullptr, "rand_cells is NULL");
    return tv_entry(links, size, data, storage, rows, data_dim, data_stride, grad_stride,
                    rand_cells, n_cells, 0, mask_out, start_dim, end_dim, scale, ignore_edge,
                    ignore_last_z, grad_data, stream);
}

extern "C" int svox2_tv_grad_sparse_run(const int32_t* links, const int32_t size[3],
                                        const void* data, int32_t storage, int64_t rows,
                                        int32_t data_dim, int64_t data_stride,
                                        int64_t grad_stride, int64_t first_cell,
                                        int64_t n_cells, uint8_t* mask_out, int32_t start_dim,
                                        int32_t end_dim, float scale, int32_t ignore_edge,
                                        int32_t ignore_last_z, float* grad_data, void* stream) {
    return tv_entry(links, size, data, storage, rows, data_dim, data_stride, grad_stride,
                    nullptr, n_cells, first_cell, mask_out, start_dim, end_dim, scale,
                    ignore_edge, ignore_last_z, grad_data, stream);
}

// ---------------------------------------------------------------------------
// Lumisphere total variation (lumisphere_tv_grad_sparse_kernel,
// svox2/csrc/loss_kernel.cu:336-446, and its launcher :643-678).  The SH
// coefficients of a cell are turned into one colour per channel by the basis
// values of a direction, A(v) = sum_k v[c B + k] sv[k], and of a perturbed
// direction, U(v) = sum_k v[c B + k] svu[k]; per listed cell and channel
//   dx, dy, dz = (A(v100 / v010 / v001) - A(v000)) * size_axis / 256,
//   du = (U(v000) - A(v000)) * dir_factor,
//   idelta = (scale / n_cells) / sqrt(1e-9 + dx^2 + dy^2 + dz^2 + du^2),
// and coefficient k of channel c receives
//   v000: (-dx sv - dy sv - dz sv + du (svu - sv)) s idelta, v001: dz sv s idelta,
//   v010: dy sv s idelta, v100: dx sv s idelta   (s = the same axis factor again),
// the gradient of sqrt(1e-9 + dx^2 + dy^2 + dz^2 + du^2) scaled by scale / n_cells.
// The reference's rules are kept: the flat cell index is split by the
// (size - 1) extents; a cell whose link is 0 is skipped; a missing v000 reads
// 0 and a missing neighbour reads v000; a term is added (and its row marked)
// where its link addresses a row and its value is nonzero; only the first 32
// coefficients of a row take part (the reference gives each cell one 32-lane
// warp, one lane per coefficient), so a channel cut by that limit sums the
// coefficients it has.  One rule is added: a flat index whose x part falls on
// or past the last slab (x >= size0 - 1; the reference's cell list spans all
// size0 * size1 * size2 indices) would read past the links array in the
// reference and is skipped here.
//
// Work layout as for the several-channel TV above: wave 0 of a workgroup reads
// 64 listed cells' links and lists the live ones in LDS; the workgroup's 8
// half-waves take them in turn, lane = coefficient.  The per-channel sums are
// segmented inclusive scans over the half-wave (5 shuffle steps) read back at
// the channel's last lane.
namespace svox2amd {
namespace {

constexpr int kLumiLanes = 32;

template <typename DT>
struct LumiJob {
    const int32_t* links;
    int s0, s1, s2;
    FastDiv div_z;  // by s2 - 1
    FastDiv div_y;  // by s1 - 1
    const DT* data;
    int rows;
    int64_t dst, gst;
    const int32_t* cells;
    int64_t n_cells;
    int dim;    // coefficients taking part: min(sh_data_dim, 32)
    int basis;  // basis_dim
    const float* sv;
    const float* svu;
    float scale;  // already divided by n_cells
    float sc0, sc1, sc2, su;
    uint8_t* mask;
    float* grad;
};

template <typename DT>
__device__ __forceinline__ bool lumi_row(const LumiJob<DT>& J, int l) {
    return static_cast<uint32_t>(l) < static_cast<uint32_t>(J.rows);
}

template <typename DT>
__device__ __forceinline__ CellLinks lumi_cell(const LumiJob<DT>& J, int64_t k) {
    CellLinks c{-1, -1, -1, -1, false};
    if (k >= J.n_cells) return c;
    const int xyz = J.cells[k];
    if (xyz < 0) return c;
    uint32_t z, y;
    const uint32_t xy = J.div_z.divmod(static_cast<uint32_t>(xyz), z);
    const uint32_t x = J.div_y.divmod(xy, y);
    if (x >= static_cast<uint32_t>(J.s0 - 1)) return c;
    const int32_t* lp = J.links + (static_cast<int64_t>(x) * J.s1 + y) * J.s2 + z;
    c.l000 = lp[0];
    if (c.l000 == 0) return c;
    c.l001 = lp[1];
    c.l010 = lp[J.s2];
    c.l100 = lp[static_cast<int64_t>(J.s1) * J.s2];
    c.live = lumi_row(J, c.l000) || lumi_row(J, c.l001) || lumi_row(J, c.l010) ||
             lumi_row(J, c.l100);
    return c;
}

// Sum of v over the lanes [c0, last] of this lane's 32-lane half (c0 <= hl <=
// last): a segmented Hillis-Steele scan, then the segment's last lane.
__device__ __forceinline__ float channel_sum(float v, int hl, int c0, int last) {
#pragma unroll
    for (int d = 1; d < kLumiLanes; d <<= 1) {
        const float o = __shfl_up(v, d, kLumiLanes);
        if (hl - d >= c0) v += o;
    }
    return __shfl(v, last, kLumiLanes);
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_tv_lumisphere(LumiJob<DT> J) {
    __shared__ int4 live[kTvWindow];
    __shared__ int n_live;
    const int hw = threadIdx.x >> 5, hl = threadIdx.x & 31;
    // This lane's coefficient, its channel's lanes and basis values.
    const bool mine = hl < J.dim;
    const int c0 = (hl / J.basis) * J.basis;
    const int last = min(c0 + J.basis, J.dim) - 1;
    const int kb = hl - c0;
    const float sv = mine ? J.sv[kb] : 0.f;
    const float svu = mine ? J.svu[kb] : 0.f;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kTvWindow; base < J.n_cells;
         base += static_cast<int64_t>(gridDim.x) * kTvWindow) {
        if (threadIdx.x < 64) {
            const CellLinks c = lumi_cell(J, base + threadIdx.x);
            const uint64_t m = __ballot(c.live);
            if (c.live)
                live[__popcll(m & ((uint64_t{1} << threadIdx.x) - 1))] =
                    make_int4(c.l000, c.l001, c.l010, c.l100);
            if (threadIdx.x == 0) n_live = static_cast<int>(__popcll(m));
        }
        __syncthreads();
        const int total = n_live;
        for (int j = hw; j < total; j += kBlock / 32) {
            const int4 L = live[j];
            const int ls[4] = {L.x, L.y, L.z, L.w};
            const bool ok000 = mine && lumi_row(J, L.x);
            const float v000 = ok000 ? ld(J.data, L.x * J.dst + hl) : 0.f;
            const float v001 = mine && lumi_row(J, L.y) ? ld(J.data, L.y * J.dst + hl) : v000;
            const float v010 = mine && lumi_row(J, L.z) ? ld(J.data, L.z * J.dst + hl) : v000;
            const float v100 = mine && lumi_row(J, L.w) ? ld(J.data, L.w * J.dst + hl) : v000;
            // every lane of the half takes part in the scans (idle lanes add 0)
            const float a000 = channel_sum(v000 * sv, hl, c0, last);
            const float a001 = channel_sum(v001 * sv, hl, c0, last);
            const float a010 = channel_sum(v010 * sv, hl, c0, last);
            const float a100 = channel_sum(v100 * sv, hl, c0, last);
            const float u000 = channel_sum(v000 * svu, hl, c0, last);
            float dx = (a100 - a000) * J.sc0;
            float dy = (a010 - a000) * J.sc1;
            float dz = (a001 - a000) * J.sc2;
            float du = (u000 - a000) * J.su;
            const float idelta = J.scale * rsqrtf(1e-9f + dx * dx + dy * dy + dz * dz + du * du);
            dx *= J.sc0;
            dy *= J.sc1;
            dz *= J.sc2;
            du *= J.su;
            const float val[4] = {-dx * sv - dy * sv - dz * sv + du * (svu - sv), dz * sv,
                                  dy * sv, dx * sv};
            bool any[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                any[i] = mine && lumi_row(J, ls[i]) && val[i] != 0.f;
                if (any[i]) atomic_add_f32(J.grad + ls[i] * J.gst + hl, val[i] * idelta);
            }
            if (J.mask) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint64_t b = __ballot(any[i]);
                    const uint32_t half = static_cast<uint32_t>(b >> (32 * (hw & 1)));
                    if (half != 0 && hl == __builtin_ctz(half)) J.mask[ls[i]] = 1;
                }
            }
        }
        __syncthreads();  // the list is rewritten by the next window
    }
}

template <typename DT>
int run_lumisphere(const svox2_grid_t& g, const DT* data, const int32_t* cells, int64_t n_cells,
                   const float* sv, const float* svu, float scale, float dir_factor,
                   const svox2_grads_t& gr, hipStream_t s) {
    LumiJob<DT> J;
    J.links = g.links;
    J.s0 = g.size[0];
    J.s1 = g.size[1];
    J.s2 = g.size[2];
    J.div_z = FastDiv(static_cast<uint32_t>(g.size[2] - 1));
    J.div_y = FastDiv(static_cast<uint32_t>(g.size[1] - 1));
    J.data = data;
    J.rows = static_cast<int>(g.rows);
    J.dst = g.sh_stride > 0 ? g.sh_stride : g.sh_data_dim;
    J.gst = gr.grad_sh_stride > 0 ? gr.grad_sh_stride : g.sh_data_dim;
    J.cells = cells;
    J.n_cells = n_cells;
    J.dim = std::min(g.sh_data_dim, kLumiLanes);
    J.basis = g.basis_dim;
    J.sv = sv;
    J.svu = svu;
    // The reference divides the scale by the number of cells (loss_kernel.cu:671).
    J.scale = scale / static_cast<float>(n_cells);
    J.sc0 = static_cast<float>(g.size[0]) * (1.f / 256.f);
    J.sc1 = static_cast<float>(g.size[1]) * (1.f / 256.f);
    J.sc2 = static_cast<float>(g.size[2]) * (1.f / 256.f);
    J.su = dir_factor;
    J.mask = gr.mask_out;
    J.grad = gr.grad_sh_out;
    const int64_t windows = (n_cells + kTvWindow - 1) / kTvWindow;
    const int64_t blocks = std::min<int64_t>(windows, 256 * 32);
    k_tv_lumisphere<DT><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(J);
    return check_launch("lumisphere_tv_grad_sparse");
}

}  // namespace
}  // namespace svox2amd

extern "C" int svox2_lumisphere_tv_grad_sparse(const svox2_grid_t* grid,
                                               const int32_t* rand_cells, int64_t n_cells,
                                               const float* basis_fn, const float* basis_fn_u,
                                               float scale, float ndc_coeffx, float ndc_coeffy,
                                               float dir_factor, const svox2_grads_t* grads,
                                               void* stream) {
    (void)ndc_coeffx;  // unused by the reference's axis factors (loss_kernel.cu:21-54)
    (void)ndc_coeffy;
    SVOX2_REQUIRE(grid != nullptr && grads != nullptr, "grid / grads are NULL");
    const svox2_grid_t& g = *grid;
    SVOX2_REQUIRE(g.links != nullptr, "links is NULL");
    for (int i = 0; i < 3; ++i) SVOX2_REQUIRE(g.size[i] >= 1, "grid size[%d] = %d", i, g.size[i]);
    SVOX2_REQUIRE(static_cast<int64_t>(g.size[0]) * g.size[1] * g.size[2] < INT32_MAX,
                  "too many cells");
    SVOX2_REQUIRE(g.storage == SVOX2_STORAGE_F32 || g.storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", g.storage);
    SVOX2_REQUIRE(g.rows >= 0 && g.rows < INT32_MAX, "rows out of range");
    SVOX2_REQUIRE(g.basis_dim >= 1 && g.sh_data_dim >= g.basis_dim,
                  "basis_dim %d invalid for sh_data_dim %d", g.basis_dim, g.sh_data_dim);
    SVOX2_REQUIRE(n_cells >= 0, "negative cell count");
    SVOX2_REQUIRE(g.sh_stride >= 0 && grads->grad_sh_stride >= 0, "negative row stride");
    SVOX2_REQUIRE(g.sh_stride == 0 || g.sh_stride >= g.sh_data_dim, "sh row stride too small");
    SVOX2_REQUIRE(grads->grad_sh_stride == 0 || grads->grad_sh_stride >= g.sh_data_dim,
                  "grad_sh row stride too small");
    // A unit extent leaves no cells between grid points.
    if (n_cells == 0 || g.rows == 0 || g.size[0] < 2 || g.size[1] < 2 || g.size[2] < 2)
        return ok();
    SVOX2_REQUIRE(rand_cells != nullptr, "rand_cells is NULL");
    SVOX2_REQUIRE(g.sh_data != nullptr && grads->grad_sh_out != nullptr,
                  "sh_data / grad_sh_out are NULL");
    SVOX2_REQUIRE(basis_fn != nullptr && basis_fn_u != nullptr, "basis values are NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (g.storage == SVOX2_STORAGE_BF16)
        return run_lumisphere(g, static_cast<const uint16_t*>(g.sh_data), rand_cells, n_cells,
                              basis_fn, basis_fn_u, scale, dir_factor, *grads, s);
    return run_lumisphere(g, static_cast<const float*>(g.sh_data), rand_cells, n_cells,
                          basis_fn, basis_fn_u, scale, dir_factor, *grads, s);
}
