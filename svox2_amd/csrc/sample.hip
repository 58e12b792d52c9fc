// Trilinear grid sampling and its adjoint for gfx950.
//
// With colours: one 32-lane group per point; lane c < C interpolates SH
// channel c and lane C interpolates density, so the eight corner rows are read
// as contiguous 108-byte segments and the adjoint issues coalesced no-return
// atomics.  Density only: one lane per point.  A single launch on the caller's
// stream replaces the reference's two kernels on two freshly created streams
// followed by host synchronisation (svox2/csrc/svox2_kernel.cu:181-267).
//
// Semantics: sample_grid_sh_kernel / sample_grid_density_kernel
// (svox2_kernel.cu:10-78) and their backward kernels (:81-176): clamp the grid
// point to [0, size-1], lower corner min(trunc, size-2), empty corners read 0.
#include "common.hpp"

#include <climits>
#include <type_traits>

namespace svox2amd {
namespace {

constexpr int kGroup = 32;
constexpr int kBlock = 256;

template <typename DT>
__device__ __forceinline__ int point_cell(const GridView<DT>& g, const float* pts, int64_t i,
                                          float* p) {
    int l[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float x = fmaf(pts[3 * i + a], g.scaling[a], g.offset[a]);
        x = fminf(fmaxf(x, 0.f), static_cast<float>(g.size[a]) - 1.f);
        l[a] = min(static_cast<int>(x), g.size[a] - 2);
        p[a] = x - static_cast<float>(l[a]);
    }
    return l[0] * g.stride_x + l[1] * g.stride_y + l[2];
}

template <typename DT>
__device__ __forceinline__ void corner_links(const GridView<DT>& g, int base, int* lk) {
    const int sx = g.stride_x, sy = g.stride_y;
    lk[0] = g.links[base];
    lk[1] = g.links[base + 1];
    lk[2] = g.links[base + sy];
    lk[3] = g.links[base + sy + 1];
    lk[4] = g.links[base + sx];
    lk[5] = g.links[base + sx + 1];
    lk[6] = g.links[base + sx + sy];
    lk[7] = g.links[base + sx + sy + 1];
}

__device__ __forceinline__ float trilerp8(const float* v, const float* p) {
    const float ix0y0 = lerpf(v[0], v[1], p[2]);
    const float ix0y1 = lerpf(v[2], v[3], p[2]);
    const float ix0 = lerpf(ix0y0, ix0y1, p[1]);
    const float ix1y0 = lerpf(v[4], v[5], p[2]);
    const float ix1y1 = lerpf(v[6], v[7], p[2]);
    const float ix1 = lerpf(ix1y0, ix1y1, p[1]);
    return lerpf(ix0, ix1, p[0]);
}

// Corner weights in the reference's product order (svox2_kernel.cu:110-128).
__device__ __forceinline__ void corner_weights(const float* p, float go, float* w) {
    const float xb = p[0], yb = p[1], zb = p[2];
    const float xa = 1.f - xb, ya = 1.f - yb, za = 1.f - zb;
    const float xago = xa * go;
    float tmp = ya * xago;
    w[0] = tmp * za;
    w[1] = tmp * zb;
    tmp = yb * xago;
    w[2] = tmp * za;
    w[3] = tmp * zb;
    const float xbgo = xb * go;
    tmp = ya * xbgo;
    w[4] = tmp * za;
    w[5] = tmp * zb;
    tmp = yb * xbgo;
    w[6] = tmp * za;
    w[7] = tmp * zb;
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_sample_all(GridView<DT> g,
                                                       const float* __restrict__ pts, int64_t n,
                                                       float* __restrict__ out_d,
                                                       float* __restrict__ out_sh) {
    const int64_t i = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    const int C = g.sh_dim;
    if (i >= n || lane > C) return;
    float p[3];
    const int base = point_cell(g, pts, i, p);
    int lk[8];
    corner_links(g, base, lk);
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const bool ok = g.valid(lk[k]);
        v[k] = !ok ? 0.f
                   : (lane < C ? ld(g.sh, static_cast<int64_t>(lk[k]) * g.sstride + lane)
                               : ld(g.density, static_cast<int64_t>(lk[k]) * g.dstride));
    }
    const float r = trilerp8(v, p);
    if (lane < C) out_sh[i * C + lane] = r;
    else out_d[i] = r;
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_sample_density(GridView<DT> g,
                                                           const float* __restrict__ pts,
                                                           int64_t n, float* __restrict__ out_d) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    float p[3];
    const int base = point_cell(g, pts, i, p);
    int lk[8];
    corner_links(g, base, lk);
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        v[k] = g.valid(lk[k]) ? ld(g.density, static_cast<int64_t>(lk[k]) * g.dstride) : 0.f;
    out_d[i] = trilerp8(v, p);
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_sample_all_backward(
        GridView<DT> g, const float* __restrict__ pts, int64_t n,
        const float* __restrict__ gout_d, const float* __restrict__ gout_sh,
        float* __restrict__ grad_d, float* __restrict__ grad_sh, int64_t gds, int64_t gss) {
    const int64_t i = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    const int C = g.sh_dim;
    if (i >= n || lane > C) return;
    float p[3];
    const int base = point_cell(g, pts, i, p);
    int lk[8];
    corner_links(g, base, lk);
    const float go = lane < C ? gout_sh[i * C + lane] : gout_d[i];
    float w[8];
    corner_weights(p, go, w);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!g.valid(lk[k])) continue;
        if (lane < C) atomic_add_f32(grad_sh + static_cast<int64_t>(lk[k]) * gss + lane, w[k]);
        else atomic_add_f32(grad_d + static_cast<int64_t>(lk[k]) * gds, w[k]);
    }
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_sample_density_backward(
        GridView<DT> g, const float* __restrict__ pts, int64_t n,
        const float* __restrict__ gout_d, float* __restrict__ grad_d, int64_t gds) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    float p[3];
    const int base = point_cell(g, pts, i, p);
    int lk[8];
    corner_links(g, base, lk);
    float w[8];
    corner_weights(p, gout_d[i], w);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (g.valid(lk[k])) atomic_add_f32(grad_d + static_cast<int64_t>(lk[k]) * gds, w[k]);
}

// Resampling (SparseGrid.resample, svox2.py:1223-1424).  The new grid's cell
// (ix, iy, iz) sits at the old grid's coordinates (ax[ix], ay[iy], az[iz]):
// the caller passes the three axis arrays, so the points are those of its own
// linspace.  Grid coordinates map to themselves (sample(grid_coords=True)).
struct NewGrid {
    const float* ax;
    const float* ay;
    const float* az;
    int n[3];
};

template <typename DT>
__device__ __forceinline__ int new_point_cell(const GridView<DT>& g, const NewGrid& ng,
                                              int64_t cell, float* p) {
    const int iz = static_cast<int>(cell % ng.n[2]);
    const int64_t ixy = cell / ng.n[2];
    const int iy = static_cast<int>(ixy % ng.n[1]);
    const int ix = static_cast<int>(ixy / ng.n[1]);
    const float q[3] = {ng.ax[ix], ng.ay[iy], ng.az[iz]};
    int l[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float x = fminf(fmaxf(q[a], 0.f), static_cast<float>(g.size[a]) - 1.f);
        l[a] = min(static_cast<int>(x), g.size[a] - 2);
        p[a] = x - static_cast<float>(l[a]);
    }
    return l[0] * g.stride_x + l[1] * g.stride_y + l[2];
}

// Pass 1: density of every new cell, one lane per cell, raster order.
template <typename DT>
__global__ __launch_bounds__(kBlock) void k_resample_density(GridView<DT> g, NewGrid ng,
                                                             int64_t n,
                                                             float* __restrict__ out) {
    const int64_t cell = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (cell >= n) return;
    float p[3];
    const int base = new_point_cell(g, ng, cell, p);
    int lk[8];
    corner_links(g, base, lk);
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        v[k] = g.valid(lk[k]) ? ld(g.density, static_cast<int64_t>(lk[k]) * g.dstride) : 0.f;
    out[cell] = trilerp8(v, p);
}

// Pass 2: the rows of the kept cells.  One 32-lane group per new row: lane
// c < C samples SH channel c at the cell's point, lane C copies the cell's
// pass-1 density; the row is written in the output's storage type.
template <typename DT, typename OT>
__global__ __launch_bounds__(kBlock) void k_resample_rows(
        GridView<DT> g, NewGrid ng, const int32_t* __restrict__ kept, int64_t n_rows,
        const float* __restrict__ density_vol, OT* __restrict__ out_sh, int64_t sh_stride,
        OT* __restrict__ out_density, int64_t density_stride) {
    const int64_t row = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    const int C = g.sh_dim;
    if (row >= n_rows || lane > C) return;
    const int64_t cell = kept[row];
    if (cell < 0 || cell >= static_cast<int64_t>(ng.n[0]) * ng.n[1] * ng.n[2]) return;
    if (lane == C) {
        st(out_density, row * density_stride, density_vol[cell]);
        return;
    }
    float p[3];
    const int base = new_point_cell(g, ng, cell, p);
    int lk[8];
    corner_links(g, base, lk);
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        v[k] = g.valid(lk[k]) ? ld(g.sh, static_cast<int64_t>(lk[k]) * g.sstride + lane) : 0.f;
    st(out_sh, row * sh_stride + lane, trilerp8(v, p));
}

template <typename Fn>
void with_view(const svox2_grid_t* grid, Fn&& fn) {
    if (grid->storage == SVOX2_STORAGE_BF16) fn(make_view<uint16_t>(*grid));
    else fn(make_view<float>(*grid));
}

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

extern "C" int svox2_sample_grid(const svox2_grid_t* grid, const float* points,
                                 int64_t n_points, int32_t want_colors, float* density_out,
                                 float* sh_out, void* stream) {
    int st = check_grid(grid, want_colors != 0);
    if (st) return st;
    SVOX2_REQUIRE(n_points >= 0, "negative point count");
    if (n_points == 0) return ok();
    SVOX2_REQUIRE(points != nullptr && density_out != nullptr, "points / density_out are NULL");
    SVOX2_REQUIRE(!want_colors || sh_out != nullptr, "sh_out is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    with_view(grid, [&](auto g) {
        using DT = typename ViewDT<decltype(g)>::type;
        if (want_colors) {
            k_sample_all<DT><<<blocks_for(n_points * kGroup, kBlock), kBlock, 0, s>>>(
                g, points, n_points, density_out, sh_out);
        } else {
            k_sample_density<DT><<<blocks_for(n_points, kBlock), kBlock, 0, s>>>(
                g, points, n_points, density_out);
        }
    });
    return check_launch("sample_grid");
}

extern "C" int svox2_sample_grid_backward(const svox2_grid_t* grid, const float* points,
                                          int64_t n_points, const float* grad_out_density,
                                          const float* grad_out_sh, float* grad_density_out,
                                          float* grad_sh_out, int64_t grad_density_stride,
                                          int64_t grad_sh_stride, int32_t want_colors,
                                          void* stream) {
    int st = check_grid(grid, false);
    if (st) return st;
    SVOX2_REQUIRE(n_points >= 0, "negative point count");
    if (n_points == 0) return ok();
    SVOX2_REQUIRE(points != nullptr && grad_out_density != nullptr && grad_density_out != nullptr,
                  "points / grad_out_density / grad_density_out are NULL");
    SVOX2_REQUIRE(!want_colors || (grad_out_sh != nullptr && grad_sh_out != nullptr),
                  "grad_out_sh / grad_sh_out are NULL");
    SVOX2_REQUIRE(grad_density_stride >= 0 && grad_sh_stride >= 0, "negative gradient stride");
    const int64_t gds = grad_density_stride > 0 ? grad_density_stride : 1;
    const int64_t gss = grad_sh_stride > 0 ? grad_sh_stride : grid->sh_data_dim;
    SVOX2_REQUIRE(gss >= grid->sh_data_dim, "grad_sh_stride below sh_data_dim");
    hipStream_t s = static_cast<hipStream_t>(stream);
    with_view(grid, [&](auto g) {
        using DT = typename ViewDT<decltype(g)>::type;
        if (want_colors) {
            k_sample_all_backward<DT><<<blocks_for(n_points * kGroup, kBlock), kBlock, 0, s>>>(
                g, points, n_points, grad_out_density, grad_out_sh, grad_density_out,
                grad_sh_out, gds, gss);
        } else {
            k_sample_density_backward<DT><<<blocks_for(n_points, kBlock), kBlock, 0, s>>>(
                g, points, n_points, grad_out_density, grad_density_out, gds);
        }
    });
    return check_launch("sample_grid_backward");
}

static int check_new_grid(const float* axes, const int32_t reso[3], NewGrid& ng, int64_t& n) {
    SVOX2_REQUIRE(axes != nullptr && reso != nullptr, "axes / reso are NULL");
    n = 1;
    for (int i = 0; i < 3; ++i) {
        SVOX2_REQUIRE(reso[i] >= 1, "reso[%d] = %d", i, reso[i]);
        ng.n[i] = reso[i];
        n *= reso[i];
    }
    SVOX2_REQUIRE(n < INT32_MAX, "new grid has too many cells");
    ng.ax = axes;
    ng.ay = axes + reso[0];
    ng.az = axes + reso[0] + reso[1];
    return ok();
}

extern "C" int svox2_resample_density(const svox2_grid_t* grid, const float* axes,
                                      const int32_t reso[3], float* density_out,
                                      void* stream) {
    int st = check_grid(grid, false);
    if (st) return st;
    NewGrid ng;
    int64_t n;
    if ((st = check_new_grid(axes, reso, ng, n))) return st;
    SVOX2_REQUIRE(density_out != nullptr, "density_out is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    with_view(grid, [&](auto g) {
        using DT = typename ViewDT<decltype(g)>::type;
        k_resample_density<DT><<<blocks_for(n, kBlock), kBlock, 0, s>>>(g, ng, n, density_out);
    });
    return check_launch("resample_density");
}

extern "C" int svox2_resample_rows(const svox2_grid_t* grid, const float* axes,
                                   const int32_t reso[3], const int32_t* kept_cells,
                                   int64_t n_rows, const float* density_volume, void* sh_out,
                                   int64_t sh_stride, void* density_out, int64_t density_stride,
                                   int32_t out_storage, void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    NewGrid ng;
    int64_t n;
    if ((st = check_new_grid(axes, reso, ng, n))) return st;
    SVOX2_REQUIRE(n_rows >= 0 && n_rows <= n, "row count %lld outside [0, %lld]",
                  (long long)n_rows, (long long)n);
    if (n_rows == 0) return ok();
    SVOX2_REQUIRE(kept_cells != nullptr && density_volume != nullptr && sh_out != nullptr &&
                      density_out != nullptr,
                  "resample_rows: NULL argument");
    SVOX2_REQUIRE(sh_stride >= grid->sh_data_dim && density_stride >= 1, "bad output strides");
    SVOX2_REQUIRE(out_storage == SVOX2_STORAGE_F32 || out_storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", out_storage);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned blocks = blocks_for(n_rows * kGroup, kBlock);
    with_view(grid, [&](auto g) {
        using DT = typename ViewDT<decltype(g)>::type;
        if (out_storage == SVOX2_STORAGE_BF16)
            k_resample_rows<DT, uint16_t><<<blocks, kBlock, 0, s>>>(
                g, ng, kept_cells, n_rows, density_volume, static_cast<uint16_t*>(sh_out),
                sh_stride, static_cast<uint16_t*>(density_out), density_stride);
        else
            k_resample_rows<DT, float><<<blocks, kBlock, 0, s>>>(
                g, ng, kept_cells, n_rows, density_volume, static_cast<float*>(sh_out),
                sh_stride, static_cast<float*>(density_out), density_stride);
    });
    return check_launch("resample_rows");
}
