// Trilinear sparse-voxel ray march (forward, backward, fused MSE, image, stats)
// for gfx950.
//
// Work mapping: one 32-lane half of a 64-wide wavefront per ray, two rays per
// wavefront, lane c holding SH channel c (27 of 32 lanes busy for SH-2).  All
// per-ray scalar state (sample position, cell, density, transmittance) is
// computed redundantly and identically by the 32 lanes, so control flow stays
// uniform inside a half; the only cross-lane traffic is the per-colour sum of
// the 9 basis-weighted channels.
//
// The eight corners of the current cell (links, densities and this lane's SH
// channel) are held in registers and re-used by every sample that falls in the
// same cell; the backward pass accumulates each corner's gradient in registers
// and issues one no-return atomic per (corner, channel) when the ray leaves the
// cell, instead of one per sample.
//
// Semantics follow the reference kernels: trace_ray_cuvol
// (svox2/csrc/render_lerp_kernel_cuvol.cu:30-120), trace_ray_cuvol_backward
// (:227-383), render_ray_kernel / render_ray_image_kernel /
// render_ray_backward_kernel (:611-773), the fused launcher (:1070-1152), and
// the helpers of svox2/csrc/include/render_util.cuh (ray_find_bounds :578-628,
// compute_skip_dist :234-295, trilerp_cuvol_one :71-91, calc_sh :332-363,
// cam2world_ray :526-544, world2ndc :504-524).
#include "common.hpp"

#include <math.h>
#include <limits.h>
#include <type_traits>

namespace svox2amd {
namespace {

constexpr int kGroup = 32;    // lanes per ray
constexpr int kBlock = 256;   // threads per workgroup
constexpr int kMaxIters = 1 << 22;  // hard per-ray iteration cap (termination guard)

struct OptView {
    float bg;
    float step;
    float sigma_thresh;
    float log_stop;  // stop when log transmittance < log_stop
    float near_clip;
    int spheric;
    int last_opaque;
};

struct Ray {
    float o[3];    // grid-space origin
    float d[3];    // unit grid-space direction
    float inv[3];  // (float)(1.0 / (double)d[i])
    float ws;      // world_step: world distance per step
    float tmin, tmax;
};

// Real SH basis function j (degree <= 2) at unit direction (x, y, z);
// constants of render_util.cuh:299-318 / svox2/utils.py:82-92.
__device__ __forceinline__ float sh_basis(int j, float x, float y, float z) {
    const float C0 = 0.28209479177387814f;
    const float C1 = 0.4886025119029199f;
    float r = C0;
    r = (j == 1) ? -C1 * y : r;
    r = (j == 2) ? C1 * z : r;
    r = (j == 3) ? -C1 * x : r;
    r = (j == 4) ? 1.0925484305920792f * (x * y) : r;
    r = (j == 5) ? -1.0925484305920792f * (y * z) : r;
    r = (j == 6) ? 0.31539156525252005f * (2.f * (z * z) - (x * x) - (y * y)) : r;
    r = (j == 7) ? -1.0925484305920792f * (x * z) : r;
    r = (j == 8) ? 0.5462742152960396f * ((x * x) - (y * y)) : r;
    return r;
}

__device__ __forceinline__ float dot3(const float* a, const float* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// World ray -> grid ray with [tmin, tmax]; ray_find_bounds (render_util.cuh:578-628)
// plus an optional per-ray (t_near, t_far) clip in world distance.
template <typename DT>
__device__ __forceinline__ void setup_ray(const GridView<DT>& g, const OptView& o,
                                          const float* wo, const float* wd,
                                          const float* clip, Ray& r) {
    float d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        r.o[i] = fmaf(wo[i], g.scaling[i], g.offset[i]);
        d[i] = wd[i] * g.scaling[i];
    }
    const float delta_scale = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        r.d[i] = d[i] * delta_scale;
        r.inv[i] = static_cast<float>(1.0 / static_cast<double>(r.d[i]));
    }
    r.ws = delta_scale * o.step;
    if (o.spheric) {
        float so[3], sd[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float s = 2.f / static_cast<float>(g.size[i]);
            so[i] = fmaf(r.o[i] + 0.5f, s, -1.f);
            sd[i] = r.d[i] * s;
        }
        const float q2a = 2.f * dot3(sd, sd);
        const float qb = 2.f * dot3(so, sd);
        const float f = qb * qb - 2.f * q2a * dot3(so, so);
        const float det_far = f + 2.f * q2a * 1.f * 1.f;
        const float rn = 1.f - o.near_clip;
        const float det_near = f + 2.f * q2a * rn * rn;
        if (det_far < 0.f || det_near < 0.f) {
            r.tmin = 1e-9f;
            r.tmax = 0.f;
        } else {
            r.tmax = (-qb + sqrtf(det_far)) / q2a;
            r.tmin = (-qb - sqrtf(det_near)) / q2a;
        }
    } else {
        r.tmin = o.near_clip / r.ws * o.step;
        r.tmax = 2e3f;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float t1 = (-0.5f - r.o[i]) * r.inv[i];
            const float t2 = (static_cast<float>(g.size[i]) - 0.5f - r.o[i]) * r.inv[i];
            if (r.d[i] != 0.f) {
                r.tmin = fmaxf(r.tmin, fminf(t1, t2));
                r.tmax = fminf(r.tmax, fmaxf(t1, t2));
            }
        }
    }
    if (clip != nullptr) {
        r.tmin = fmaxf(r.tmin, clip[0] / r.ws * o.step);
        r.tmax = fminf(r.tmax, clip[1] / r.ws * o.step);
    }
}

// Distance to the exit of the empty 2^s block around cell l, 0 if not worth
// skipping; compute_skip_dist (render_util.cuh:234-295).
__device__ __forceinline__ float skip_dist(int link, const int* l, const float* p,
                                           const Ray& r) {
    if (link >= -1) return 0.f;
    const uint32_t shift = static_cast<uint32_t>(-static_cast<int64_t>(link)) - 1u;
    const uint32_t side = static_cast<uint32_t>(static_cast<float>(1 << shift) - 1.f);
    float tmin = 0.f, tmax = 1e9f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int ul = ((l[i] >> shift) << shift) - l[i];
        const float t1 = (static_cast<float>(ul) - p[i]) * r.inv[i];
        const float t2 = (static_cast<float>(static_cast<uint32_t>(ul) + side) - p[i]) * r.inv[i];
        if (r.d[i] != 0.f) {
            tmin = fmaxf(tmin, fminf(t1, t2));
            tmax = fminf(tmax, fmaxf(t1, t2));
        }
    }
    return (tmin > 0.f) ? 0.f : tmax;
}

// Trilinear interpolation of eight corner values in the reference order:
// z, then y, then x (trilerp_cuvol_one, render_util.cuh:71-91).  Corner k has
// bits (dx << 2) | (dy << 1) | dz.
__device__ __forceinline__ float trilerp8(const float* v, const float* p) {
    const float ix0y0 = lerpf(v[0], v[1], p[2]);
    const float ix0y1 = lerpf(v[2], v[3], p[2]);
    const float ix0 = lerpf(ix0y0, ix0y1, p[1]);
    const float ix1y0 = lerpf(v[4], v[5], p[2]);
    const float ix1y1 = lerpf(v[6], v[7], p[2]);
    const float ix1 = lerpf(ix1y0, ix1y1, p[1]);
    return lerpf(ix0, ix1, p[0]);
}

// Sums each colour's B basis-weighted channels inside the 32-lane group and
// broadcasts the three sums to every lane.  j = lane % B.
template <int B>
__device__ __forceinline__ void color_sums(float v, int j, float* col) {
    if (B > 1) { const float u = __shfl_down(v, 1, kGroup); if (j + 1 < B) v += u; }
    if (B > 2) { const float u = __shfl_down(v, 2, kGroup); if (j + 2 < B) v += u; }
    if (B > 4) { const float u = __shfl_down(v, 4, kGroup); if (j + 4 < B) v += u; }
    if (B > 8) { const float u = __shfl_down(v, 8, kGroup); if (j + 8 < B) v += u; }
    col[0] = __shfl(v, 0, kGroup);
    col[1] = __shfl(v, B, kGroup);
    col[2] = __shfl(v, 2 * B, kGroup);
}

__device__ __forceinline__ float pick3(const float* a, int i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}

struct Fnv {
    uint32_t h = 2166136261u;
    __device__ __forceinline__ void mix(uint32_t w) { h = (h ^ w) * 16777619u; }
};

// Per-ray lane context.
template <int B>
struct Lane {
    int lane;    // 0..31
    int j;       // basis index lane % B
    int color;   // lane / B (clamped to 2)
    bool act;    // lane < 3B
    int ch;      // channel loaded by this lane
    float basis; // SH basis j at the view direction
    __device__ __forceinline__ Lane(int lane_, const float* vdir) {
        lane = lane_;
        act = lane < 3 * B;
        j = lane % B;
        color = act ? lane / B : 2;
        ch = act ? lane : 0;
        basis = act ? sh_basis(j, vdir[0], vdir[1], vdir[2]) : 0.f;
    }
};

// The register window: links, densities and this lane's SH channel at the
// eight corners of cell (c0, c1, c2).
template <typename DT, int B>
struct Window {
    int c0 = INT_MIN, c1 = 0, c2 = 0;
    int lk[8];
    float dn[8];
    float sv[8];
    bool full = false;  // links 1..7 and densities loaded
    bool shv = false;   // SH channel loaded

    __device__ __forceinline__ bool enter(const GridView<DT>& g, const int* l, int base) {
        if (l[0] == c0 && l[1] == c1 && l[2] == c2) return false;
        c0 = l[0]; c1 = l[1]; c2 = l[2];
        lk[0] = g.links[base];
        full = false;
        shv = false;
        return true;
    }
    __device__ __forceinline__ void load_density(const GridView<DT>& g, int base) {
        if (full) return;
        const int sx = g.stride_x, sy = g.stride_y;
        lk[1] = g.links[base + 1];
        lk[2] = g.links[base + sy];
        lk[3] = g.links[base + sy + 1];
        lk[4] = g.links[base + sx];
        lk[5] = g.links[base + sx + 1];
        lk[6] = g.links[base + sx + sy];
        lk[7] = g.links[base + sx + sy + 1];
#pragma unroll
        for (int k = 0; k < 8; ++k) dn[k] = g.valid(lk[k]) ? ld(g.density, lk[k]) : 0.f;
        full = true;
    }
    __device__ __forceinline__ void load_sh(const GridView<DT>& g, int ch) {
        if (shv) return;
        const int C = 3 * B;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            sv[k] = g.valid(lk[k]) ? ld(g.sh, static_cast<int64_t>(lk[k]) * C + ch) : 0.f;
        shv = true;
    }
};

// Sample position of step t: clamp to the grid, lower cell corner l and the
// fractional offset p (the loop head of trace_ray_cuvol, :55-62).
template <typename DT>
__device__ __forceinline__ int sample_cell(const GridView<DT>& g, const Ray& r, float t,
                                           float* p, int* l) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float x = fmaf(t, r.d[i], r.o[i]);
        x = fminf(fmaxf(x, 0.f), static_cast<float>(g.size[i]) - 1.f);
        l[i] = min(static_cast<int>(x), g.size[i] - 2);
        p[i] = x - static_cast<float>(l[i]);
    }
    return l[0] * g.stride_x + l[1] * g.stride_y + l[2];
}

// Forward march of one ray.  out[3] = composited colour (all lanes),
// returns the final log transmittance.
template <typename DT, int B, bool STATS>
__device__ __forceinline__ float march_forward(const GridView<DT>& g, const OptView& o,
                                               const Ray& r, const Lane<B>& ln,
                                               float* out, int32_t* stats) {
    out[0] = out[1] = out[2] = 0.f;
    uint32_t n_iter = 0, n_eval = 0, n_hit = 0;
    Fnv fnv;
    float log_t = 0.f;
    if (r.tmin <= r.tmax) {
        Window<DT, B> w;
        float t = r.tmin;
        float ws = r.ws;
        int guard = 0;
        while (t <= r.tmax && ++guard < kMaxIters) {
            if (STATS) ++n_iter;
            float p[3];
            int l[3];
            const int base = sample_cell(g, r, t, p, l);
            w.enter(g, l, base);
            if (w.lk[0] < -1) {
                const float skip = skip_dist(w.lk[0], l, p, r);
                if (skip >= o.step) {
                    t += ceilf(skip / o.step) * o.step;
                    continue;
                }
            }
            if (STATS) {
                ++n_eval;
                fnv.mix(static_cast<uint32_t>(base));
                fnv.mix(__float_as_uint(t));
            }
            w.load_density(g, base);
            const float sigma = trilerp8(w.dn, p);
            if (o.last_opaque && t + o.step > r.tmax) ws = 1e9f;
            if (sigma > o.sigma_thresh) {
                if (STATS) { ++n_hit; fnv.mix(1u); }
                w.load_sh(g, ln.ch);
                const float v = ln.act ? trilerp8(w.sv, p) * ln.basis : 0.f;
                float col[3];
                color_sums<B>(v, ln.j, col);
                const float pcnt = ws * sigma;
                const float weight = __expf(log_t) * (1.f - __expf(-pcnt));
                log_t -= pcnt;
#pragma unroll
                for (int c = 0; c < 3; ++c) out[c] += weight * fmaxf(col[c] + 0.5f, 0.f);
                if (log_t < o.log_stop) {
                    log_t = -1e3f;
                    break;
                }
            } else if (STATS) {
                fnv.mix(0u);
            }
            t += o.step;
        }
    }
    const float tb = __expf(log_t) * o.bg;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] += tb;
    if (STATS && ln.lane == 0) {
        fnv.mix(__float_as_uint(log_t));
        stats[0] = static_cast<int32_t>(n_iter);
        stats[1] = static_cast<int32_t>(n_eval);
        stats[2] = static_cast<int32_t>(n_hit);
        stats[3] = static_cast<int32_t>(fnv.h);
    }
    return log_t;
}

struct GradOut {
    float* __restrict__ gd;   // grad density (rows)
    float* __restrict__ gs;   // grad sh (rows x C)
    uint8_t* __restrict__ mask;
};

// Backward march of one ray given dL/drgb (gq) and accum = rgb . gq (+ beta
// term); accumulates into the gradient buffers (trace_ray_cuvol_backward,
// render_lerp_kernel_cuvol.cu:227-383).
template <typename DT, int B>
__device__ __forceinline__ void march_backward(const GridView<DT>& g, const OptView& o,
                                               const Ray& r, const Lane<B>& ln,
                                               const float* gq, float accum,
                                               float sparsity, const GradOut& go) {
    if (r.tmin > r.tmax) return;
    const int C = 3 * B;
    const float gout = pick3(gq, ln.color);
    Window<DT, B> w;
    float acc_c[8], acc_d[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { acc_c[k] = 0.f; acc_d[k] = 0.f; }
    bool touched = false;

    auto flush = [&]() {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int lk = w.lk[k];
            if (g.valid(lk)) {
                if (go.gs != nullptr && ln.act)
                    atomic_add_f32(go.gs + static_cast<int64_t>(lk) * C + ln.lane,
                                   ln.basis * acc_c[k]);
                if (ln.lane == 0) {
                    if (go.gd != nullptr) atomic_add_f32(go.gd + lk, acc_d[k]);
                    if (go.mask != nullptr) go.mask[lk] = 1;
                }
            }
            acc_c[k] = 0.f;
            acc_d[k] = 0.f;
        }
    };

    float t = r.tmin;
    float ws = r.ws;
    float log_t = 0.f;
    int guard = 0;
    while (t <= r.tmax && ++guard < kMaxIters) {
        float p[3];
        int l[3];
        const int base = sample_cell(g, r, t, p, l);
        if (!(l[0] == w.c0 && l[1] == w.c1 && l[2] == w.c2)) {
            if (touched) { flush(); touched = false; }
            w.enter(g, l, base);
        }
        if (w.lk[0] < -1) {
            const float skip = skip_dist(w.lk[0], l, p, r);
            if (skip >= o.step) {
                t += ceilf(skip / o.step) * o.step;
                continue;
            }
        }
        w.load_density(g, base);
        const float sigma = trilerp8(w.dn, p);
        if (o.last_opaque && t + o.step > r.tmax) ws = 1e9f;
        if (sigma > o.sigma_thresh) {
            w.load_sh(g, ln.ch);
            const float v = ln.act ? trilerp8(w.sv, p) * ln.basis : 0.f;
            float col[3];
            color_sums<B>(v, ln.j, col);
            const float pcnt = ws * sigma;
            const float weight = __expf(log_t) * (1.f - __expf(-pcnt));
            log_t -= pcnt;
            float total = 0.f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                col[c] += 0.5f;
                total += fmaxf(col[c], 0.f) * gq[c];
            }
            const float mine = pick3(col, ln.color);
            const float in01 = (fmaxf(mine, 0.f) == mine) ? 1.f : 0.f;
            const float grad_common = weight * in01 * gout;
            accum -= weight * total;
            float grad_sigma = ws * (total * __expf(log_t) - accum);
            if (sparsity > 0.f)
                grad_sigma += sparsity * (4.f * sigma / (1.f + 2.f * (sigma * sigma)));
            // Trilinear weights of the eight corners.
            const float ax = 1.f - p[0], ay = 1.f - p[1], az = 1.f - p[2];
            const float wyz[4] = {ay * az, ay * p[2], p[1] * az, p[1] * p[2]};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float wk = wyz[k & 3] * ((k & 4) ? p[0] : ax);
                acc_c[k] = fmaf(wk, grad_common, acc_c[k]);
                acc_d[k] = fmaf(wk, grad_sigma, acc_d[k]);
            }
            touched = true;
            if (log_t < o.log_stop) break;
        }
        t += o.step;
    }
    if (touched) flush();
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <typename DT, int B, bool STATS>
__global__ __launch_bounds__(kBlock) void k_render_rays(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        float* __restrict__ rgb_out, float* __restrict__ log_t_out,
        int32_t* __restrict__ stats) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (ray >= n) return;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const Lane<B> ln(lane, wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    float out[3];
    const float log_t = march_forward<DT, B, STATS>(g, o, r, ln, out,
                                                    STATS ? stats + 4 * ray : nullptr);
    if (!STATS) {
        if (lane < 3) rgb_out[3 * ray + lane] = pick3(out, lane);
        if (log_t_out != nullptr && lane == 0) log_t_out[ray] = log_t;
    }
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

template <typename DT, int B>
__global__ __launch_bounds__(kBlock) void k_render_image(
        GridView<DT> g, OptView o, svox2_camera_t cam, float* __restrict__ rgb_out,
        float* __restrict__ log_t_out) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    const int64_t n = static_cast<int64_t>(cam.width) * cam.height;
    if (ray >= n) return;
    const int ix = static_cast<int>(ray % cam.width);
    const int iy = static_cast<int>(ray / cam.width);
    float wd[3], wo[3];
    camera_ray(cam, ix, iy, wd, wo);
    const Lane<B> ln(lane, wd);
    Ray r;
    setup_ray(g, o, wo, wd, nullptr, r);
    float out[3];
    const float log_t = march_forward<DT, B, false>(g, o, r, ln, out, nullptr);
    if (lane < 3) rgb_out[3 * ray + lane] = pick3(out, lane);
    if (log_t_out != nullptr && lane == 0) log_t_out[ray] = log_t;
}

template <typename DT, int B>
__global__ __launch_bounds__(kBlock) void k_render_backward(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        const float* __restrict__ grad_out, const float* __restrict__ color_cache,
        GradOut go) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (ray >= n) return;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const Lane<B> ln(lane, wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    const float gq[3] = {grad_out[3 * ray], grad_out[3 * ray + 1], grad_out[3 * ray + 2]};
    const float* cc = color_cache + 3 * ray;
    const float accum = fmaf(cc[0], gq[0], fmaf(cc[1], gq[1], cc[2] * gq[2]));
    march_backward<DT, B>(g, o, r, ln, gq, accum, 0.f, go);
}

// Forward then backward of the same ray in one pass of the workgroup: the
// forward colour stays in registers and feeds the MSE gradient
// g = (rgb - gt) * 2 / (3n) (render_ray_backward_kernel, :735-741).
template <typename DT, int B>
__global__ __launch_bounds__(kBlock) void k_render_fused(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        const float* __restrict__ rgb_gt, float beta_per_ray, float sparsity,
        float norm, float* __restrict__ rgb_out, GradOut go) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (ray >= n) return;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const Lane<B> ln(lane, wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    float rgb[3];
    const float log_t = march_forward<DT, B, false>(g, o, r, ln, rgb, nullptr);
    if (lane < 3) rgb_out[3 * ray + lane] = pick3(rgb, lane);
    float gq[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gq[c] = (rgb[c] - rgb_gt[3 * ray + c]) * norm;
    float accum = fmaf(rgb[0], gq[0], fmaf(rgb[1], gq[1], rgb[2] * gq[2]));
    if (beta_per_ray > 0.f) {
        const float tin = __expf(log_t);
        accum += beta_per_ray * (1.f - tin / (1.f - tin + 1e-3f));
    }
    march_backward<DT, B>(g, o, r, ln, gq, accum, sparsity, go);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

float log_stop_threshold(float stop_thresh) {
    // exp(x) < stop  <=>  x < ln(stop); the float bound is rounded up so that
    // the float comparison x < bound equals the real one.
    if (!(stop_thresh > 0.f)) return -INFINITY;
    const double L = log(static_cast<double>(stop_thresh));
    float f = static_cast<float>(L);
    if (static_cast<double>(f) < L) f = nextafterf(f, INFINITY);
    return f;
}

int make_opt(const svox2_render_options_t* opt, OptView& o) {
    SVOX2_REQUIRE(opt != nullptr, "render options are NULL");
    SVOX2_REQUIRE(opt->step_size >= 1e-3f && opt->step_size <= 1e3f,
                  "step_size must be in [1e-3, 1e3] voxels, got %g", (double)opt->step_size);
    o.bg = opt->background_brightness;
    o.step = opt->step_size;
    o.sigma_thresh = opt->sigma_thresh;
    o.log_stop = log_stop_threshold(opt->stop_thresh);
    o.near_clip = opt->near_clip;
    o.spheric = opt->use_spheric_clip != 0;
    o.last_opaque = opt->last_sample_opaque != 0;
    return ok();
}

int check_rays(const float* origins, const float* dirs, int64_t n) {
    SVOX2_REQUIRE(n >= 0, "negative ray count %lld", (long long)n);
    SVOX2_REQUIRE(n == 0 || (origins != nullptr && dirs != nullptr), "ray origins/dirs are NULL");
    SVOX2_REQUIRE(n < (INT64_C(1) << 40), "too many rays");
    return ok();
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

template <typename Fn>
int dispatch(const svox2_grid_t* grid, Fn&& fn) {
    const int bd = grid->basis_dim;
    const bool bf = grid->storage == SVOX2_STORAGE_BF16;
#define SVOX2_CASE(B)                                                        \
    if (bd == B) {                                                           \
        if (bf) fn(make_view<uint16_t>(*grid), std::integral_constant<int, B>(), (uint16_t*)nullptr); \
        else fn(make_view<float>(*grid), std::integral_constant<int, B>(), (float*)nullptr);       \
        return SVOX2_OK;                                                     \
    }
    SVOX2_CASE(1)
    SVOX2_CASE(4)
    SVOX2_CASE(9)
#undef SVOX2_CASE
    return fail(SVOX2_ERR_UNSUPPORTED, "basis_dim %d not supported (1, 4 or 9)", bd);
}

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

extern "C" int svox2_volume_render_cuvol(const svox2_grid_t* grid, const float* origins,
                                         const float* dirs, const float* ray_clip,
                                         int64_t n_rays, const svox2_render_options_t* opt,
                                         float* rgb_out, float* log_transmit_out,
                                         void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_out != nullptr, "rgb_out is NULL");
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_rays<DT, B, false>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, origins, dirs, ray_clip, n_rays, rgb_out,
                           log_transmit_out, (int32_t*)nullptr);
    });
    if (st) return st;
    return check_launch("volume_render_cuvol");
}

extern "C" int svox2_volume_render_cuvol_stats(const svox2_grid_t* grid, const float* origins,
                                               const float* dirs, const float* ray_clip,
                                               int64_t n_rays, const svox2_render_options_t* opt,
                                               int32_t* stats_out, void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(stats_out != nullptr, "stats_out is NULL");
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_rays<DT, B, true>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, origins, dirs, ray_clip, n_rays,
                           (float*)nullptr, (float*)nullptr, stats_out);
    });
    if (st) return st;
    return check_launch("volume_render_cuvol_stats");
}

extern "C" int svox2_volume_render_cuvol_image(const svox2_grid_t* grid,
                                               const svox2_camera_t* cam,
                                               const svox2_render_options_t* opt,
                                               float* rgb_out, float* log_transmit_out,
                                               void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    SVOX2_REQUIRE(cam != nullptr, "camera is NULL");
    SVOX2_REQUIRE(cam->width >= 0 && cam->height >= 0, "negative image size");
    const int64_t n = static_cast<int64_t>(cam->width) * cam->height;
    if (n == 0) return ok();
    SVOX2_REQUIRE(rgb_out != nullptr, "rgb_out is NULL");
    const unsigned blocks = blocks_for(n * kGroup, kBlock);
    const svox2_camera_t c = *cam;
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_image<DT, B>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, c, rgb_out, log_transmit_out);
    });
    if (st) return st;
    return check_launch("volume_render_cuvol_image");
}

extern "C" int svox2_volume_render_cuvol_backward(const svox2_grid_t* grid,
                                                  const float* origins, const float* dirs,
                                                  const float* ray_clip, int64_t n_rays,
                                                  const svox2_render_options_t* opt,
                                                  const float* grad_out,
                                                  const float* color_cache,
                                                  const svox2_grads_t* grads,
                                                  void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    SVOX2_REQUIRE(grads != nullptr, "grads is NULL");
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(grad_out != nullptr && color_cache != nullptr,
                  "grad_out / color_cache are NULL");
    const GradOut go{grads->grad_density_out, grads->grad_sh_out, grads->mask_out};
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_backward<DT, B>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, origins, dirs, ray_clip, n_rays, grad_out,
                           color_cache, go);
    });
    if (st) return st;
    return check_launch("volume_render_cuvol_backward");
}

extern "C" int svox2_volume_render_cuvol_fused(const svox2_grid_t* grid,
                                               const float* origins, const float* dirs,
                                               const float* ray_clip, int64_t n_rays,
                                               const svox2_render_options_t* opt,
                                               const float* rgb_gt, float beta_loss,
                                               float sparsity_loss, float* rgb_out,
                                               const svox2_grads_t* grads, void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    SVOX2_REQUIRE(grads != nullptr, "grads is NULL");
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_gt != nullptr && rgb_out != nullptr, "rgb_gt / rgb_out are NULL");
    const GradOut go{grads->grad_density_out, grads->grad_sh_out, grads->mask_out};
    const float norm = 2.f / static_cast<float>(3 * n_rays);
    const float beta = beta_loss / static_cast<float>(n_rays);
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_fused<DT, B>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, origins, dirs, ray_clip, n_rays, rgb_gt,
                           beta, sparsity_loss, norm, rgb_out, go);
    });
    if (st) return st;
    return check_launch("volume_render_cuvol_fused");
}
