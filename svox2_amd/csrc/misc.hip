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
