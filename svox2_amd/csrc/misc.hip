// Empty-space distance codes for the ray march (accel_dist_prop) on gfx950.
//
// Pass 1 marks, for every occupied cell, each ancestor of an occupancy
// pyramid (sizes halved with ceiling while all three exceed 1); pass 2
// rewrites every empty link to -1 - k, k = number of consecutive empty
// ancestors.  The pyramid lives in the caller's workspace and is cleared by a
// fill kernel on the same stream (no memset node, so the sequence can be
// captured in a graph).
//
// Semantics: accel_dist_set_kernel / accel_dist_prop_kernel
// (svox2/csrc/misc_kernel.cu:109-181) and the host function (:356-416).
#include "common.hpp"

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

__global__ __launch_bounds__(kBlock) void k_accel_set(const int32_t* __restrict__ links, int s0,
                                                      int s1, int s2,
                                                      uint8_t* __restrict__ tmp) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (tid >= static_cast<int64_t>(s0) * s1 * s2) return;
    int z = static_cast<int>(tid % s2);
    const int64_t xy = tid / s2;
    int y = static_cast<int>(xy % s1);
    int x = static_cast<int>(xy / s1);
    if (links[tid] < 0) return;
    int sx = s0, sy = s1, sz = s2;
    uint8_t* base = tmp;
    while (sx > 1 && sy > 1 && sz > 1) {
        x >>= 1;
        y >>= 1;
        z >>= 1;
        sx = half_ceil(sx);
        sy = half_ceil(sy);
        sz = half_ceil(sz);
        base[(static_cast<int64_t>(x) * sy + y) * sz + z] = 1;
        base += static_cast<int64_t>(sx) * sy * sz;
    }
}

__global__ __launch_bounds__(kBlock) void k_accel_prop(int32_t* __restrict__ links, int s0,
                                                       int s1, int s2,
                                                       const uint8_t* __restrict__ tmp) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (tid >= static_cast<int64_t>(s0) * s1 * s2) return;
    if (links[tid] >= 0) return;
    int z = static_cast<int>(tid % s2);
    const int64_t xy = tid / s2;
    int y = static_cast<int>(xy % s1);
    int x = static_cast<int>(xy / s1);
    int sx = s0, sy = s1, sz = s2;
    const uint8_t* base = tmp;
    int result = -1;
    while (sx > 1 && sy > 1 && sz > 1) {
        x >>= 1;
        y >>= 1;
        z >>= 1;
        sx = half_ceil(sx);
        sy = half_ceil(sy);
        sz = half_ceil(sz);
        if (base[(static_cast<int64_t>(x) * sy + y) * sz + z]) break;
        result -= 1;
        base += static_cast<int64_t>(sx) * sy * sz;
    }
    links[tid] = result;
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
    const unsigned blocks = blocks_for(cells, kBlock);
    k_accel_set<<<blocks, kBlock, 0, s>>>(links, size[0], size[1], size[2], tmp);
    k_accel_prop<<<blocks, kBlock, 0, s>>>(links, size[0], size[1], size[2], tmp);
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
