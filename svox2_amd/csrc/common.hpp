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

// World ray of pixel (ix, iy) (cam2world_ray, render_util.cuh:526-544; NDC
// warp world2ndc :504-524 when ndc_coeffx > 0).
__device__ __forceinline__ void camera_ray(const svox2_camera_t& cam, int ix, int iy,
                                           float* dir, float* origin) {
    float x = (static_cast<float>(ix) + 0.5f - cam.cx) / cam.fx;
    float y = (static_cast<float>(iy) + 0.5f - cam.cy) / cam.fy;
    float z = sqrtf(static_cast<float>(static_cast<double>(x * x + y * y) + 1.0));
    x /= z;
    y /= z;
    z = 1.0f / z;
    const float* m = cam.c2w;
    dir[0] = m[0] * x + m[1] * y + m[2] * z;
    dir[1] = m[4] * x + m[5] * y + m[6] * z;
    dir[2] = m[8] * x + m[9] * y + m[10] * z;
    origin[0] = m[3];
    origin[1] = m[7];
    origin[2] = m[11];
    if (cam.ndc_coeffx > 0.f) {
        const float nearp = 1.f;
        const float t = (nearp - origin[2]) / dir[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) origin[i] = fmaf(t, dir[i], origin[i]);
        dir[0] = cam.ndc_coeffx * (dir[0] / dir[2] - origin[0] / origin[2]);
        dir[1] = cam.ndc_coeffy * (dir[1] / dir[2] - origin[1] / origin[2]);
        dir[2] = 2.f * nearp / origin[2];
        origin[0] = cam.ndc_coeffx * (origin[0] / origin[2]);
        origin[1] = cam.ndc_coeffy * (origin[1] / origin[2]);
        origin[2] = 1.f - 2.f * nearp / origin[2];
        const float rn = 1.0f / sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        dir[0] *= rn;
        dir[1] *= rn;
        dir[2] *= rn;
    }
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

// MSI background of a grid (background.hip).  grid_has_background: whether
// the grid carries one.  check_background validates it with the render
// options.  launch_background adds each ray's background colour (rays from
// origins / dirs, or from cam when not NULL) to rgb behind the grid's log
// transmittance log_t_fg, writing it to bg_rgb_out too when not NULL.
// launch_background_backward accumulates the background's gradients (see
// k_background_backward for grad_in / rgb_residual and bg_src / bg_src_is_fg).
bool grid_has_background(const svox2_grid_t* g);
int check_background(const svox2_grid_t* g, const svox2_render_options_t* opt);
int launch_background(const svox2_grid_t* g, const svox2_render_options_t* opt,
                      const float* origins, const float* dirs, const svox2_camera_t* cam,
                      int64_t n, const float* log_t_fg, float* rgb, float* bg_rgb_out,
                      hipStream_t s);
int launch_background_backward(const svox2_grid_t* g, const svox2_render_options_t* opt,
                               const float* origins, const float* dirs, int64_t n,
                               const float* grad_in, const float* color_cache, int rgb_residual,
                               float norm, const float* log_t_fg, const float* bg_src,
                               int bg_src_is_fg, float sparsity, const svox2_grads_t* grads,
                               hipStream_t s);

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
