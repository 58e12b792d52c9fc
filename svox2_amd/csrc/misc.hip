// Empty-space distance codes for the ray march (accel_dist_prop) on gfx950.
//
// Pass 1 marks, for every occupied cell, each ancestor of an occupancy
// pyramid (sizes halved with ceiling while all three exceed 1); pass 2 turns
// each level-0 node's flag into the number of empty levels from it up to its
// first occupied ancestor; pass 3 rewrites every empty link to -1 - k, k =
// number of consecutive empty ancestors, from its parent node's count (one
// byte read per cell instead of a walk up the pyramid).  The pyramid lives in
// the caller's workspace and is cleared by a fill kernel on the same stream
// (no memset node, so the sequence can be captured in a graph).
//
// Semantics: accel_dist_set_kernel / accel_dist_prop_kernel
// (svox2/csrc/misc_kernel.cu:109-181) and the host function (:356-416).
#include "common.hpp"
#include "fastdiv.hpp"

namespace svox2amd {
namespace {

constexpr int kBlock = 256;

__host__ __device__ __forceinline__ int half_ceil(int x) { return (x + 1) / 2; }

int64_t pyramid_bytes(int sx, int sy, int sz) {
    int64_t req = 0;
    while (sx > 1 && sy > 1 && sz > 1) {
        sx = half_ceil(sx);
        sy = half_ceil(sy);
        sz = half_ceil(sz);
        req += static_cast<int64_t>(sx) * sy * sz;
    }
    return req;
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(uint32_t* __restrict__ p, int64_t n_words) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n_words) p[i] = 0u;
}

// The occupancy pyramid's levels: level k (k = 0 .. n - 1) has size
// (.., sy[k], sz[k]) and starts at byte off[k] of the workspace.  Cells
// number below 2^31, so the smallest axis is below 1291 and there are at
// most 11 levels.
constexpr int kMaxLevels = 16;
struct Pyramid {
    int n;
    int sy[kMaxLevels];
    int sz[kMaxLevels];
    int64_t off[kMaxLevels];
};

Pyramid make_pyramid(int sx, int sy, int sz) {
    Pyramid p{};
    int64_t off = 0;
    while (sx > 1 && sy > 1 && sz > 1 && p.n < kMaxLevels) {
        sx = half_ceil(sx);
        sy = half_ceil(sy);
        sz = half_ceil(sz);
        p.sy[p.n] = sy;
        p.sz[p.n] = sz;
        p.off[p.n] = off;
        off += static_cast<int64_t>(sx) * sy * sz;
        ++p.n;
    }
    return p;
}

// The set and prop passes: workgroups run over the grid's (x, y) rows
// (blockIdx.x = x * s1 + y, one uniform division per workgroup) and
// kBlock * kCells z cells each (blockIdx.y); a thread takes kCells cells
// kBlock apart, so its loads of all of them are in flight together.
constexpr int kCells = 8;

__device__ __forceinline__ void row_cells(int s1, int& x, int& y, int& z0) {
    const int row = static_cast<int>(blockIdx.x);
    x = row / s1;
    y = row - x * s1;
    z0 = static_cast<int>(blockIdx.y) * (kBlock * kCells) + static_cast<int>(threadIdx.x);
}

// Index of the level-k ancestor (k >= 0) of cell (x, y, z) in the pyramid.
__device__ __forceinline__ int64_t ancestor(const Pyramid& p, int k, int x, int y, int z) {
    return p.off[k] + (static_cast<int64_t>(x >> (k + 1)) * p.sy[k] + (y >> (k + 1))) * p.sz[k] +
           (z >> (k + 1));
}

// Pass 1: every occupied cell marks all its ancestors.
__global__ __launch_bounds__(kBlock) void k_accel_set(const int32_t* __restrict__ links, int s1,
                                                      int s2, Pyramid pyr,
                                                      uint8_t* __restrict__ tmp) {
    int x, y, z0;
    row_cells(s1, x, y, z0);
    const int64_t row = (static_cast<int64_t>(x) * s1 + y) * s2;
    int32_t l[kCells];
#pragma unroll
    for (int j = 0; j < kCells; ++j) {
        const int z = z0 + j * kBlock;
        l[j] = z < s2 ? links[row + z] : -1;
    }
#pragma unroll
    for (int j = 0; j < kCells; ++j) {
        const int z = z0 + j * kBlock;
        if (l[j] < 0) continue;
#pragma unroll
        for (int k = 0; k < kMaxLevels; ++k)
            if (k < pyr.n) tmp[ancestor(pyr, k, x, y, z)] = 1;
    }
}

// Pass 2, over the level-0 nodes: each node's byte becomes the number of
// empty pyramid levels from it up to its first occupied ancestor (0 if it is
// occupied, n if none is).  A node's byte is read and written by its own
// thread only; the flags of levels 1 and up are read together.
__global__ __launch_bounds__(kBlock) void k_accel_node(Pyramid pyr, uint8_t* __restrict__ tmp) {
    const int row = static_cast<int>(blockIdx.x);
    const int a = row / pyr.sy[0];
    const int b = row - a * pyr.sy[0];
    const int c = static_cast<int>(blockIdx.y) * kBlock + static_cast<int>(threadIdx.x);
    if (c >= pyr.sz[0]) return;
    const int64_t self = pyr.off[0] + (static_cast<int64_t>(a) * pyr.sy[0] + b) * pyr.sz[0] + c;
    const bool own = tmp[self] != 0;
    uint32_t occ = 0u;
#pragma unroll
    for (int k = 1; k < kMaxLevels; ++k) {
        if (k < pyr.n) {
            const int64_t i = pyr.off[k] +
                              (static_cast<int64_t>(a >> k) * pyr.sy[k] + (b >> k)) * pyr.sz[k] +
                              (c >> k);
            occ |= (tmp[i] != 0 ? 1u : 0u) << k;
        }
    }
    tmp[self] = static_cast<uint8_t>(own ? 0 : (occ ? __ffs(static_cast<int>(occ)) - 1 : pyr.n));
}

// Pass 3: every empty link becomes -1 - (its level-0 parent's byte): -1 - k
// for k consecutive empty ancestors (-1 with no pyramid).
__global__ __launch_bounds__(kBlock) void k_accel_prop(int32_t* __restrict__ links, int s1,
                                                       int s2, Pyramid pyr,
                                                       const uint8_t* __restrict__ tmp) {
    int x, y, z0;
    row_cells(s1, x, y, z0);
    const int64_t row = (static_cast<int64_t>(x) * s1 + y) * s2;
    int32_t l[kCells];
    int e[kCells];
#pragma unroll
    for (int j = 0; j < kCells; ++j) {
        const int z = z0 + j * kBlock;
        l[j] = z < s2 ? links[row + z] : 0;
        e[j] = (z < s2 && pyr.n > 0) ? tmp[ancestor(pyr, 0, x, y, z)] : 0;
    }
#pragma unroll
    for (int j = 0; j < kCells; ++j) {
        const int z = z0 + j * kBlock;
        if (z < s2 && l[j] < 0) links[row + z] = -1 - e[j];
    }
}

// Vector forms of passes 1 and 3 for grids whose rows (z) hold a multiple of
// 4 cells: a thread takes 4 consecutive cells of one row with one 16-byte
// load (and, in pass 3, one 16-byte store; occupied links are stored back
// unchanged), kVec such groups kBlock * 4 cells apart over the flat cell
// index.  The 4 cells' level-0 parents are 2 adjacent bytes, read as one
// 16-bit word.  The streaming passes then move 16 bytes per lane.
constexpr int kVec = 4;

// Splits of the flat cell index (< 2^31) by the row length s2 and by s1, as
// invariant-divisor multiplies.
struct CellDiv {
    FastDiv s2, s1;
};

__device__ __forceinline__ void cell_xyz(uint32_t cell, const CellDiv& cd, int& x, int& y,
                                         int& z) {
    uint32_t zr, yr;
    const uint32_t row = cd.s2.divmod(cell, zr);
    x = static_cast<int>(cd.s1.divmod(row, yr));
    y = static_cast<int>(yr);
    z = static_cast<int>(zr);
}

__global__ __launch_bounds__(kBlock) void k_accel_set4(const int4* __restrict__ links, int64_t n4,
                                                       CellDiv cd, Pyramid pyr,
                                                       uint8_t* __restrict__ tmp) {
    const int64_t v0 = static_cast<int64_t>(blockIdx.x) * (kBlock * kVec) + threadIdx.x;
    int4 l[kVec];
#pragma unroll
    for (int j = 0; j < kVec; ++j) {
        const int64_t v = v0 + j * kBlock;
        l[j] = v < n4 ? links[v] : make_int4(-1, -1, -1, -1);
    }
#pragma unroll
    for (int j = 0; j < kVec; ++j) {
        const int c[4] = {l[j].x, l[j].y, l[j].z, l[j].w};
        if (max(max(c[0], c[1]), max(c[2], c[3])) < 0) continue;
        int x, y, z;
        cell_xyz(static_cast<uint32_t>(4 * (v0 + j * kBlock)), cd, x, y, z);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (c[q] < 0) continue;
#pragma unroll
            for (int k = 0; k < kMaxLevels; ++k)
                if (k < pyr.n) tmp[ancestor(pyr, k, x, y, z + q)] = 1;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_accel_prop4(int4* __restrict__ links, int64_t n4,
                                                        CellDiv cd, Pyramid pyr,
                                                        const uint8_t* __restrict__ tmp) {
    const int64_t v0 = static_cast<int64_t>(blockIdx.x) * (kBlock * kVec) + threadIdx.x;
    int4 l[kVec];
    uint32_t e[kVec];
#pragma unroll
    for (int j = 0; j < kVec; ++j) {
        const int64_t v = v0 + j * kBlock;
        l[j] = make_int4(0, 0, 0, 0);
        e[j] = 0u;
        if (v < n4) {
            l[j] = links[v];
            if (pyr.n > 0) {
                int x, y, z;
                cell_xyz(static_cast<uint32_t>(4 * v), cd, x, y, z);
                e[j] = *reinterpret_cast<const uint16_t*>(tmp + ancestor(pyr, 0, x, y, z));
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kVec; ++j) {
        const int64_t v = v0 + j * kBlock;
        if (v >= n4) continue;
        const int e0 = -1 - static_cast<int>(e[j] & 0xFFu);  // parent of cells 0 and 1
        const int e1 = -1 - static_cast<int>(e[j] >> 8);     // parent of cells 2 and 3
        int4 out;
        out.x = l[j].x < 0 ? e0 : l[j].x;
        out.y = l[j].y < 0 ? e0 : l[j].y;
        out.z = l[j].z < 0 ? e1 : l[j].z;
        out.w = l[j].w < 0 ? e1 : l[j].w;
        links[v] = out;
    }
}

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

extern "C" size_t svox2_accel_dist_prop_workspace_bytes(const int32_t size[3]) {
    if (size == nullptr || size[0] < 1 || size[1] < 1 || size[2] < 1) return 0;
    // Rounded up to whole 4-byte words for the clearing kernel.
    const int64_t b = pyramid_bytes(size[0], size[1], size[2]);
    return static_cast<size_t>((b + 3) / 4 * 4);
}

extern "C" int svox2_accel_dist_prop(int32_t* links, const int32_t size[3], void* workspace,
                                     size_t workspace_bytes, void* stream) {
    SVOX2_REQUIRE(links != nullptr && size != nullptr, "links / size are NULL");
    for (int i = 0; i < 3; ++i) SVOX2_REQUIRE(size[i] >= 1, "grid size[%d] = %d", i, size[i]);
    const int64_t cells = static_cast<int64_t>(size[0]) * size[1] * size[2];
    SVOX2_REQUIRE(cells < INT32_MAX, "too many cells");
    const size_t need = svox2_accel_dist_prop_workspace_bytes(size);
    SVOX2_REQUIRE(workspace_bytes >= need, "workspace has %zu bytes, %zu needed",
                  workspace_bytes, need);
    SVOX2_REQUIRE(need == 0 || workspace != nullptr, "workspace is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t* tmp = static_cast<uint8_t*>(workspace);
    if (need > 0) {
        const int64_t words = static_cast<int64_t>(need / 4);
        k_fill_zero<<<blocks_for(words, kBlock), kBlock, 0, s>>>(
            static_cast<uint32_t*>(workspace), words);
    }
    const Pyramid pyr = make_pyramid(size[0], size[1], size[2]);
    // Fewer than 2^31 cells give at most 11 levels.
    SVOX2_REQUIRE(pyr.n < kMaxLevels, "grid too large for the occupancy pyramid");
    SVOX2_REQUIRE(size[2] <= (1 << 24), "grid size[2] = %d above 2^24", size[2]);
    // Rows of a multiple of 4 cells (and 16-byte aligned links) take the
    // vector passes; the level-0 parents of 4 such cells then sit at an even
    // byte offset of the pyramid.
    const bool vec = size[2] % 4 == 0 && reinterpret_cast<uintptr_t>(links) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(tmp) % 2 == 0;
    const dim3 grid(static_cast<unsigned>(static_cast<int64_t>(size[0]) * size[1]),
                    blocks_for(size[2], kBlock * kCells));
    const int64_t n4 = cells / 4;
    const CellDiv cd{FastDiv(static_cast<uint32_t>(size[2])),
                     FastDiv(static_cast<uint32_t>(size[1]))};
    const unsigned vblocks = blocks_for(n4, kBlock * kVec);
    if (vec)
        k_accel_set4<<<vblocks, kBlock, 0, s>>>(reinterpret_cast<const int4*>(links), n4, cd, pyr,
                                                tmp);
    else
        k_accel_set<<<grid, kBlock, 0, s>>>(links, size[1], size[2], pyr, tmp);
    if (pyr.n > 0) {
        const dim3 nodes(static_cast<unsigned>(static_cast<int64_t>(half_ceil(size[0])) *
                                               pyr.sy[0]),
                         blocks_for(pyr.sz[0], kBlock));
        k_accel_node<<<nodes, kBlock, 0, s>>>(pyr, tmp);
    }
    if (vec)
        k_accel_prop4<<<vblocks, kBlock, 0, s>>>(reinterpret_cast<int4*>(links), n4, cd, pyr, tmp);
    else
        k_accel_prop<<<grid, kBlock, 0, s>>>(links, size[1], size[2], pyr, tmp);
    return check_launch("accel_dist_prop");
}

// ---------------------------------------------------------------------------
// Cell map: per 2x2x2 block of cells, the shared skip level of its lower
// corners and one occupancy bit per cell (any corner link addresses a row).
// One thread per block reads the up to 3x3x3 links the block's cells touch.
// ---------------------------------------------------------------------------
namespace svox2amd {
namespace {

__global__ __launch_bounds__(kBlock) void k_build_cell_map(const int32_t* __restrict__ links,
                                                           int s0, int s1, int s2, int rows,
                                                           int m0, int m1, int m2,
                                                           uint16_t* __restrict__ map) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (tid >= static_cast<int64_t>(m0) * m1 * m2) return;
    const int bz = static_cast<int>(tid % m2);
    const int by = static_cast<int>((tid / m2) % m1);
    const int bx = static_cast<int>(tid / (static_cast<int64_t>(m1) * m2));
    // v[i][j][k] = link at (2bx+i, 2by+j, 2bz+k), or -1 past the grid.
    int v[3][3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int x = 2 * bx + i, y = 2 * by + j, z = 2 * bz + k;
                v[i][j][k] = (x < s0 && y < s1 && z < s2)
                                 ? links[(static_cast<int64_t>(x) * s1 + y) * s2 + z]
                                 : -1;
            }
    int level = -1;  // -1: none seen yet
    bool mixed = false;
    uint32_t occ = 0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                // Cell with lower corner (2bx+i, 2by+j, 2bz+k) exists iff that
                // corner is at most size-2 on every axis.
                if (2 * bx + i > s0 - 2 || 2 * by + j > s1 - 2 || 2 * bz + k > s2 - 2) continue;
                const int l0 = v[i][j][k];
                const int lv = l0 < -1 ? -l0 - 1 : 0;
                if (level < 0) level = lv;
                else if (level != lv) mixed = true;
                bool any = false;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int lk = v[i + (c >> 2)][j + ((c >> 1) & 1)][k + (c & 1)];
                    any |= static_cast<uint32_t>(lk) < static_cast<uint32_t>(rows);
                }
                if (any) occ |= 1u << (4 * i + 2 * j + k);
            }
    if (level < 0) level = 0;
    const uint32_t lvl = (mixed || level > 254) ? 255u : static_cast<uint32_t>(level);
    map[tid] = static_cast<uint16_t>(lvl | (occ << 8));
}

}  // namespace
}  // namespace svox2amd

extern "C" size_t svox2_cell_map_bytes(const int32_t size[3]) {
    if (size == nullptr || size[0] < 2 || size[1] < 2 || size[2] < 2) return 0;
    return static_cast<size_t>(cell_map_blocks(size[0])) * cell_map_blocks(size[1]) *
           cell_map_blocks(size[2]) * sizeof(uint16_t);
}

extern "C" int svox2_build_cell_map(const int32_t* links, const int32_t size[3], int64_t rows,
                                    uint16_t* map_out, size_t map_bytes, void* stream) {
    SVOX2_REQUIRE(links != nullptr && size != nullptr && map_out != nullptr,
                  "links / size / map_out are NULL");
    for (int i = 0; i < 3; ++i) SVOX2_REQUIRE(size[i] >= 2, "grid size[%d] = %d", i, size[i]);
    const int64_t cells = static_cast<int64_t>(size[0]) * size[1] * size[2];
    SVOX2_REQUIRE(cells < INT32_MAX, "too many cells");
    SVOX2_REQUIRE(rows >= 0 && rows < INT32_MAX, "rows %lld out of range", (long long)rows);
    const size_t need = svox2_cell_map_bytes(size);
    SVOX2_REQUIRE(map_bytes >= need, "map has %zu bytes, %zu needed", map_bytes, need);
    const int m0 = cell_map_blocks(size[0]), m1 = cell_map_blocks(size[1]),
              m2 = cell_map_blocks(size[2]);
    const int64_t blocks = static_cast<int64_t>(m0) * m1 * m2;
    k_build_cell_map<<<blocks_for(blocks, kBlock), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
        links, size[0], size[1], size[2], static_cast<int>(rows), m0, m1, m2, map_out);
    return check_launch("build_cell_map");
}
