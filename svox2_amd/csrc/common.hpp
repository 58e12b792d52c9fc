// Shared device helpers and host-side argument checking for the gfx950
// sparse-voxel kernels.  Every kernel here is built with -ffp-contract=off so
// that the index arithmetic (sample positions, cell indices, skip distances,
// trilinear density) rounds exactly as written and can be reproduced bit for
// bit by a plain C restatement; fused multiply-adds are spelled out with fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include <string>

#include "../../include/svox2_amd.h"

namespace svox2amd {

// ---------------------------------------------------------------------------
// Host-side error reporting (thread-local last error, returned as status).
// ---------------------------------------------------------------------------
std::string& last_error_slot();

inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error_slot() = buf;
    return code;
}

inline int ok() {
    last_error_slot().clear();
    return SVOX2_OK;
}

// Checks the launch that was just enqueued.
inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        return fail(SVOX2_ERR_HIP, "%s: kernel launch failed: %s", what,
                    hipGetErrorString(e));
    }
    return ok();
}

#define SVOX2_REQUIRE(cond, ...)                                   \
    do {                                                           \
        if (!(cond)) return ::svox2amd::fail(SVOX2_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
    } while (0)

// ---------------------------------------------------------------------------
// Storage element loads.  BF16 values are raw bit patterns (uint16_t); the
// widening to float is exact.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float ld(const float* p, int64_t i) { return p[i]; }
__device__ __forceinline__ float ld(const uint16_t* p, int64_t i) {
    return __uint_as_float(static_cast<uint32_t>(p[i]) << 16);
}

// Raw storage bits of element i (float bits, or bf16 bits in the low half),
// and their float value: loads behind a predicate are written as raw loads
// followed by the conversions, so that the conversions (and the waits they
// need) come after all the loads instead of inside each predicated load.
__device__ __forceinline__ uint32_t ld_bits(const float* p, int64_t i) {
    return __float_as_uint(p[i]);
}
__device__ __forceinline__ uint32_t ld_bits(const uint16_t* p, int64_t i) {
    return static_cast<uint32_t>(p[i]);
}
template <typename DT>
__device__ __forceinline__ float bits_value(uint32_t b);
template <>
__device__ __forceinline__ float bits_value<float>(uint32_t b) { return __uint_as_float(b); }
template <>
__device__ __forceinline__ float bits_value<uint16_t>(uint32_t b) { return __uint_as_float(b << 16); }

// Round-to-nearest-even float -> bf16 bits (NaN kept quiet).
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
        return static_cast<uint16_t>((u >> 16) | 0x40u);
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ void st(float* p, int64_t i, float v) { p[i] = v; }
__device__ __forceinline__ void st(uint16_t* p, int64_t i, float v) {
    p[i] = f32_to_bf16_bits(v);
}

// (1 - w) a + w b with one rounding for the product-add, the form the
// reference's trilinear interpolation uses (include/cuda_util.cuh:72-75).
__device__ __forceinline__ float lerpf(float a, float b, float w) {
    return fmaf(w, b - a, a);
}

// No-return float atomic add; lowers to global_atomic_add_f32.
__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
    unsafeAtomicAdd(p, v);
}

// ---------------------------------------------------------------------------
// Grid view passed by value to kernels.
// ---------------------------------------------------------------------------
// Cell-map blocks along an axis of n grid points (cells 0..n-2, two per block).
__host__ __device__ inline int cell_map_blocks(int n) { return (n - 2) / 2 + 1; }

template <typename DT>
struct GridView {
    const int32_t* __restrict__ links;
    const DT* __restrict__ density;
    const DT* __restrict__ sh;
    int size[3];
    int stride_x;  // size[1] * size[2]
    int stride_y;  // size[2]
    int rows;
    int basis_dim;
    int sh_dim;
    int dstride;   // elements between density rows
    int sstride;   // elements between SH rows
    bool vec_rows; // SH rows are 16-byte aligned with room for whole 16-byte loads
    float offset[3];
    float scaling[3];
    const uint16_t* __restrict__ cmap;  // cell map (may be null)
    int mstride_x;  // map blocks per x step: mb[1] * mb[2]
    int mstride_y;  // mb[2]
    // A link addresses a data row iff 0 <= link < rows; negative links are
    // empty cells and out-of-range links are treated as empty.
    __device__ __forceinline__ bool valid(int link) const {
        return static_cast<uint32_t>(link) < static_cast<uint32_t>(rows);
    }
};

template <typename V>
struct ViewDT;
template <typename DT>
struct ViewDT<GridView<DT>> {
    using type = DT;
};

template <typename DT>
__host__ __device__ inline GridView<DT> make_view(const svox2_grid_t& g) {
    GridView<DT> v;
    v.links = g.links;
    v.density = static_cast<const DT*>(g.density_data);
    v.sh = static_cast<const DT*>(g.sh_data);
    for (int i = 0; i < 3; ++i) {
        v.size[i] = g.size[i];
        v.offset[i] = g.offset[i];
        v.scaling[i] = g.scaling[i];
    }
    v.stride_x = g.size[1] * g.size[2];
    v.stride_y = g.size[2];
    v.rows = static_cast<int>(g.rows);
    v.basis_dim = g.basis_dim;
    v.sh_dim = g.sh_data_dim;
    v.dstride = g.density_stride > 0 ? static_cast<int>(g.density_stride) : 1;
    v.sstride = g.sh_stride > 0 ? static_cast<int>(g.sh_stride) : g.sh_data_dim;
    // Whole 16-byte loads of a row stay inside it when the row stride is a
    // multiple of 16 bytes (it is then >= the channel count rounded up to 16 bytes).
    const int per16 = 16 / static_cast<int>(sizeof(DT));
    v.vec_rows = (reinterpret_cast<uintptr_t>(g.sh_data) % 16 == 0) && (v.sstride % per16 == 0);
    v.cmap = g.cell_map;
    v.mstride_y = cell_map_blocks(g.size[2]);
    v.mstride_x = cell_map_blocks(g.size[1]) * v.mstride_y;
    return v;
}

// Validates a grid descriptor; `need_sh` requires sh_data to be present.
int check_grid(const svox2_grid_t* g, bool need_sh);

// Resolved row strides of a gradient descriptor.
inline int64_t grad_dstride(const svox2_grads_t& g) {
    return g.grad_density_stride > 0 ? g.grad_density_stride : 1;
}
inline int64_t grad_sstride(const svox2_grads_t& g, int C) {
    return g.grad_sh_stride > 0 ? g.grad_sh_stride : C;
}

inline unsigned blocks_for(int64_t n, int threads) {
    return static_cast<unsigned>((n + threads - 1) / threads);
}

}  // namespace svox2amd
