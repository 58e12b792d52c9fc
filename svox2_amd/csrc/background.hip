// The MSI background (multi-sphere image) composited behind the grid, its
// gradients, and its total variation.
//
// The background is nlayers concentric spheres around the grid's unit sphere,
// each an equirectangular image of [2 reso][reso] pixels holding (r, g, b,
// sigma); links map a pixel column to a row of background_data
// ([rows][nlayers] float4), -1 = empty.  A ray leaving the grid with log
// transmittance log_t (the grid's own, rendered with background brightness 0)
// samples the spheres at radii n_steps / (n_steps - i - 0.5), i = 0 ..
// n_steps - 1 (n_steps = nlayers / step_size + 2), from the inside out; each
// sample is trilinear in (longitude, latitude, layer), its step is the change
// of 1 / radius times the ray's world step, and it composites like the grid's
// samples (render_background_forward / _backward,
// render_lerp_kernel_cuvol.cu:386-607; kernels :775-860).  The transmittance
// left after the last layer sees the background brightness.
//
// One 32-lane group per ray, one lane per sample: a chunk of 32 samples is
// placed and read in parallel (8 float4 loads per sample: 4 pixel columns x 2
// layers, all four channels at once), its transmittance is an inclusive scan
// of the samples' optical depths, and the stop threshold cuts the chunk at the
// first sample whose transmittance falls below it.  The reference walks the
// samples one by one in a thread per ray; the sums here are the same terms
// added in scan order.  The backward then scatters sample by sample with a
// lane per (column, layer, channel), so that each atomic instruction covers
// 8 consecutive floats of each column.
#include "common.hpp"

namespace svox2amd {
namespace {

constexpr int kBgBlock = 256;
constexpr int kBgGroup = 32;
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kInvPi = 0.318309886183790671538f;

struct BgView {
    const int32_t* links;  // [2 reso][reso]
    const float4* data;    // [rows * nlayers]
    float* grad;           // [rows * nlayers * 4] (backward)
    uint8_t* mask;         // [rows * nlayers] (backward, may be NULL)
    int64_t rows;
    int reso, nlayers, n_steps;
    float offset[3], scaling[3], sph[3];  // sph = 2 / grid size
    float bg, stop;                       // background brightness, stop threshold
};

// The ray in the frame where the grid's sphere is the unit sphere
// (ray_find_bounds_bg, render_util.cuh:630-657), and its sphere
// intersection coefficients (ConcentricSpheresIntersector, :546-576).
struct BgRay {
    float o[3], d[3];
    float ws;     // world length per unit of (1 / radius) change
    float q2a, qb, f;
    float inner;  // first radius: max(distance of the line to the centre + 1e-3, 1)
};

__device__ __forceinline__ float bg_dot(const float* a, const float* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void bg_ray(const BgView& b, const float* wo, const float* wd,
                                       BgRay& r) {
    float o[3], d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        o[i] = fmaf(wo[i], b.scaling[i], b.offset[i]);
        d[i] = wd[i] * b.scaling[i];
    }
    float ws = 1.0f / sqrtf(bg_dot(d, d));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        d[i] *= ws;
        r.o[i] = fmaf(o[i] + 0.5f, b.sph[i], -1.f);
        r.d[i] = d[i] * b.sph[i];
    }
    const float inorm = 1.0f / sqrtf(bg_dot(r.d, r.d));
    r.ws = ws * inorm;
#pragma unroll
    for (int i = 0; i < 3; ++i) r.d[i] *= inorm;
    r.q2a = 2.f * bg_dot(r.d, r.d);
    r.qb = 2.f * bg_dot(r.o, r.d);
    r.f = r.qb * r.qb - 2.f * r.q2a * bg_dot(r.o, r.o);
    const float cx = r.o[1] * r.d[2] - r.o[2] * r.d[1];
    const float cy = r.o[2] * r.d[0] - r.o[0] * r.d[2];
    const float cz = r.o[0] * r.d[1] - r.o[1] * r.d[0];
    r.inner = fmaxf(sqrtf(cx * cx + cy * cy + cz * cz) + 1e-3f, 1.f);
}

// Sample i of the ray: where it meets the sphere of radius n_steps /
// (n_steps - i - 0.5) on the way out (skipped inside the first radius or where
// it misses), 1 / |position|, and the trilinear cell and fractions in
// (longitude, latitude, layer) (render_lerp_kernel_cuvol.cu:403-428).
struct BgSample {
    bool valid;
    float invr;
    int l[3];
    float f[3];
};

__device__ __forceinline__ BgSample bg_sample(const BgView& b, const BgRay& r, int i) {
    BgSample s;
    s.valid = false;
    s.invr = 0.f;
    if (i >= b.n_steps) return s;
    const float rad = static_cast<float>(b.n_steps) / (static_cast<float>(b.n_steps - i) - 0.5f);
    if (rad < r.inner) return s;
    const float det = r.f + 2.f * r.q2a * rad * rad;
    if (det < 0.f) return s;
    const float t = (-r.qb + sqrtf(det)) / r.q2a;
    float p[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) p[j] = fmaf(t, r.d[j], r.o[j]);
    const float invr = 1.0f / sqrtf(bg_dot(p, p));
#pragma unroll
    for (int j = 0; j < 3; ++j) p[j] *= invr;
    // Equirectangular pixel of the unit direction (_unitvec2equirect,
    // render_util.cuh:482-490).
    const float lat = asinf(p[1]);
    const float lon = atan2f(p[0], p[2]);
    const float x = static_cast<float>(2 * b.reso) * (0.5f + lon * (0.5f * kInvPi));
    const float y = static_cast<float>(b.reso) * (0.5f - lat * kInvPi);
    const float z = fminf(fmaxf((1.f - invr) * static_cast<float>(b.nlayers) - 0.5f, 0.f),
                          static_cast<float>(b.nlayers - 1));
    // Truncation toward zero as the reference's (int) casts; the lower clamp
    // only keeps a NaN position (a direction rounded just past a pole) inside.
    s.l[0] = max(min(static_cast<int>(x), 2 * b.reso - 1), 0);
    s.l[1] = max(min(static_cast<int>(y), b.reso - 1), 0);
    s.l[2] = max(min(static_cast<int>(z), b.nlayers - 2), 0);
    s.f[0] = x - static_cast<float>(s.l[0]);
    s.f[1] = y - static_cast<float>(s.l[1]);
    s.f[2] = z - static_cast<float>(s.l[2]);
    s.invr = invr;
    s.valid = true;
    return s;
}

// Row offset (link * nlayers + layer) of pixel column (x, y), or -1.
__device__ __forceinline__ int64_t bg_row(const BgView& b, int x, int y, int l2) {
    const int link = b.links[x * b.reso + y];
    if (link < 0 || link >= b.rows) return -1;
    return static_cast<int64_t>(link) * b.nlayers + l2;
}

__device__ __forceinline__ float4 lerp4(float4 a, float4 c, float w) {
    return make_float4(lerpf(a.x, c.x, w), lerpf(a.y, c.y, w), lerpf(a.z, c.z, w),
                       lerpf(a.w, c.w, w));
}

__device__ __forceinline__ float4 bg_column(const BgView& b, int64_t row, float f2) {
    if (row < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    return lerp4(b.data[row], b.data[row + 1], f2);
}

// The four pixel columns of a sample: (x, y), (x, ny), (nx, y), (nx, ny),
// wrapping at the image edges (trilerp_bg_one, render_util.cuh:157-189).
struct BgCorners {
    int64_t row[4];
};

__device__ __forceinline__ BgCorners bg_corners(const BgView& b, const BgSample& s) {
    const int ny = s.l[1] < b.reso - 1 ? s.l[1] + 1 : 0;
    const int nx = s.l[0] < 2 * b.reso - 1 ? s.l[0] + 1 : 0;
    BgCorners c;
    c.row[0] = bg_row(b, s.l[0], s.l[1], s.l[2]);
    c.row[1] = bg_row(b, s.l[0], ny, s.l[2]);
    c.row[2] = bg_row(b, nx, s.l[1], s.l[2]);
    c.row[3] = bg_row(b, nx, ny, s.l[2]);
    return c;
}

__device__ __forceinline__ float4 bg_value(const BgView& b, const BgSample& s,
                                           const BgCorners& c) {
    const float4 v00 = bg_column(b, c.row[0], s.f[2]);
    const float4 v01 = bg_column(b, c.row[1], s.f[2]);
    const float4 v10 = bg_column(b, c.row[2], s.f[2]);
    const float4 v11 = bg_column(b, c.row[3], s.f[2]);
    return lerp4(lerp4(v00, v01, s.f[1]), lerp4(v10, v11, s.f[1]), s.f[0]);
}

// 32-lane group helpers (the group is one half of the wavefront; every lane of
// the group is active wherever these are called).
__device__ __forceinline__ uint32_t bg_ballot(bool p) {
    return static_cast<uint32_t>(__ballot(p) >> (threadIdx.x & 32u));
}
__device__ __forceinline__ float bg_scan(float v, int lane) {
#pragma unroll
    for (int d = 1; d < kBgGroup; d <<= 1) {
        const float u = __shfl_up(v, d, kBgGroup);
        if (lane >= d) v += u;
    }
    return v;
}
__device__ __forceinline__ float bg_sum(float v) {
#pragma unroll
    for (int m = kBgGroup / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kBgGroup);
    return v;
}

// Per-chunk march state shared by the forward and the backward.
struct BgChunk {
    BgSample s;
    BgCorners c;
    float4 v;
    float il;        // 1 / radius of the previous placed sample
    float pcnt;      // optical depth of the sample (0 unless it contributes)
    float lt_after;  // log transmittance after the sample
    float incl;      // inclusive scan of pcnt
    bool live;       // contributes and is not past the stop
    int stop_lane;   // the chunk's last live lane or 31
    bool stopped;
};

// Places and reads chunk `base` of the ray; advances invr_last.  lt is the log
// transmittance before the chunk.
__device__ __forceinline__ void bg_chunk(const BgView& b, const BgRay& r, int base, int lane,
                                         float lt, float& invr_last, BgChunk& k) {
    k.s = bg_sample(b, r, base + lane);
    const uint32_t vm = bg_ballot(k.s.valid);
    const uint32_t below = vm & ((1u << lane) - 1u);
    const int src = below ? 31 - static_cast<int>(__clz(below)) : lane;
    const float prev = __shfl(k.s.invr, src, kBgGroup);
    k.il = below ? prev : invr_last;
    if (vm) invr_last = __shfl(k.s.invr, 31 - static_cast<int>(__clz(vm)), kBgGroup);
    k.v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k.s.valid) {
        k.c = bg_corners(b, k.s);
        k.v = bg_value(b, k.s, k.c);
    }
    const bool contrib = k.s.valid && k.v.w > 0.f;
    k.pcnt = contrib ? (k.il - k.s.invr) * r.ws * k.v.w : 0.f;
    k.incl = bg_scan(k.pcnt, lane);
    k.lt_after = lt - k.incl;
    const uint32_t stopm = bg_ballot(contrib && __expf(k.lt_after) < b.stop);
    k.stopped = stopm != 0u;
    k.stop_lane = k.stopped ? __ffs(stopm) - 1 : kBgGroup - 1;
    k.live = contrib && lane <= k.stop_lane;
}

__device__ __forceinline__ float pick3f(const float* v, int c) {
    return c == 0 ? v[0] : (c == 1 ? v[1] : v[2]);
}

// Background colour of each ray added to rgb [n][3] (render_background_kernel
// / render_background_image_kernel, render_lerp_kernel_cuvol.cu:775-818); rays
// whose grid log transmittance is below -25 get nothing.  bg_rgb_out (may be
// NULL) receives the added colour.
template <bool IMAGE>
__global__ __launch_bounds__(kBgBlock) void k_background(
        BgView b, const float* __restrict__ origins, const float* __restrict__ dirs,
        svox2_camera_t cam, int64_t n, const float* __restrict__ log_t_fg,
        float* __restrict__ rgb, float* __restrict__ bg_rgb_out) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBgBlock + threadIdx.x) / kBgGroup;
    const int lane = threadIdx.x % kBgGroup;
    if (ray >= n) return;
    float lt = log_t_fg[ray];
    if (lt < -25.f) {
        if (bg_rgb_out != nullptr && lane < 3) bg_rgb_out[3 * ray + lane] = 0.f;
        return;
    }
    float wo[3], wd[3];
    if (IMAGE) {
        camera_ray(cam, static_cast<int>(ray % cam.width), static_cast<int>(ray / cam.width), wd,
                   wo);
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            wo[i] = origins[3 * ray + i];
            wd[i] = dirs[3 * ray + i];
        }
    }
    BgRay r;
    bg_ray(b, wo, wd, r);
    float invr_last = 1.f / r.inner;
    float acc[3] = {0.f, 0.f, 0.f};
    for (int base = 0; base < b.n_steps; base += kBgGroup) {
        BgChunk k;
        bg_chunk(b, r, base, lane, lt, invr_last, k);
        if (k.live) {
            const float w = __expf(k.lt_after + k.pcnt) * (1.f - __expf(-k.pcnt));
            acc[0] += w * fmaxf(fmaf(k.v.x, kShC0, 0.5f), 0.f);
            acc[1] += w * fmaxf(fmaf(k.v.y, kShC0, 0.5f), 0.f);
            acc[2] += w * fmaxf(fmaf(k.v.z, kShC0, 0.5f), 0.f);
        }
        lt -= __shfl(k.incl, k.stop_lane, kBgGroup);
        if (k.stopped) break;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = bg_sum(acc[c]);
    if (lane < 3) {
        const float val = pick3f(acc, lane) + __expf(lt) * b.bg;
        rgb[3 * ray + lane] += val;
        if (bg_rgb_out != nullptr) bg_rgb_out[3 * ray + lane] = val;
    }
}

// Background gradients of each ray (render_background_backward_kernel,
// render_lerp_kernel_cuvol.cu:820-860).  dL/drgb is grad_in, or with
// rgb_residual (color_cache - grad_in) * norm (grad_in then holds the
// targets).  The background's own colour is bg_src, or with bg_src_is_fg
// color_cache - bg_src (bg_src then holds the grid's colours); its dot product
// with dL/drgb is the running `accum` the reference carries over from the
// grid's backward.
__global__ __launch_bounds__(kBgBlock) void k_background_backward(
        BgView b, const float* __restrict__ origins, const float* __restrict__ dirs, int64_t n,
        const float* __restrict__ grad_in, const float* __restrict__ color_cache,
        int rgb_residual, float norm, const float* __restrict__ log_t_fg,
        const float* __restrict__ bg_src, int bg_src_is_fg, float sparsity) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBgBlock + threadIdx.x) / kBgGroup;
    const int lane = threadIdx.x % kBgGroup;
    if (ray >= n) return;
    float lt = log_t_fg[ray];
    if (lt < -25.f) return;
    float gq[3], bgc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float cc = color_cache[3 * ray + c];
        gq[c] = rgb_residual ? (cc - grad_in[3 * ray + c]) * norm : grad_in[3 * ray + c];
        bgc[c] = bg_src_is_fg ? cc - bg_src[3 * ray + c] : bg_src[3 * ray + c];
    }
    float accum = fmaf(bgc[0], gq[0], fmaf(bgc[1], gq[1], bgc[2] * gq[2]));
    float wo[3], wd[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        wo[i] = origins[3 * ray + i];
        wd[i] = dirs[3 * ray + i];
    }
    BgRay r;
    bg_ray(b, wo, wd, r);
    float invr_last = 1.f / r.inner;
    for (int base = 0; base < b.n_steps; base += kBgGroup) {
        BgChunk k;
        bg_chunk(b, r, base, lane, lt, invr_last, k);
        float w = 0.f, col[3] = {0.f, 0.f, 0.f}, total = 0.f;
        if (k.live) {
            w = __expf(k.lt_after + k.pcnt) * (1.f - __expf(-k.pcnt));
            col[0] = fmaf(k.v.x, kShC0, 0.5f);
            col[1] = fmaf(k.v.y, kShC0, 0.5f);
            col[2] = fmaf(k.v.z, kShC0, 0.5f);
#pragma unroll
            for (int c = 0; c < 3; ++c) total += fmaxf(col[c], 0.f) * gq[c];
        }
        const float acc_incl = bg_scan(w * total, lane);
        // This sample's terms: per channel (colours where the colour is
        // positive, the density term always) and its trilinear weights.
        float v[4] = {0.f, 0.f, 0.f, 0.f}, wxy[4] = {0.f, 0.f, 0.f, 0.f};
        float wz0 = 0.f, wz1 = 0.f;
        int rows4[4] = {-1, -1, -1, -1};
        if (k.live) {
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = col[c] > 0.f ? kShC0 * w * gq[c] : 0.f;
            const float accum_after = accum - acc_incl;
            float gs = r.ws * (k.il - k.s.invr) * (total * __expf(k.lt_after) - accum_after);
            if (sparsity > 0.f)
                gs += sparsity * (4.f * k.v.w / (1.f + 2.f * (k.v.w * k.v.w)));
            v[3] = gs;
            const float ax = 1.f - k.s.f[0], ay = 1.f - k.s.f[1];
            wxy[0] = ay * ax;
            wxy[1] = k.s.f[1] * ax;
            wxy[2] = ay * k.s.f[0];
            wxy[3] = k.s.f[1] * k.s.f[0];
            wz0 = 1.f - k.s.f[2];
            wz1 = k.s.f[2];
#pragma unroll
            for (int q = 0; q < 4; ++q) rows4[q] = static_cast<int>(k.c.row[q]);
        }
        // Scatter sample by sample, lane = (column, layer, channel): one
        // atomic instruction per sample whose lanes add into 8 consecutive
        // floats (2 layers x 4 channels) of each of the 4 columns, so they
        // coalesce into a few line requests (a lane per sample issuing its
        // own 32 atomics made each one a request of its own).
        const int q = lane >> 3, layer = (lane >> 2) & 1, ch = lane & 3;
        uint32_t todo = bg_ballot(k.live);
        while (todo != 0u) {
            const int j = __ffs(todo) - 1;
            todo &= todo - 1u;
            const float vj[4] = {__shfl(v[0], j, kBgGroup), __shfl(v[1], j, kBgGroup),
                                 __shfl(v[2], j, kBgGroup), __shfl(v[3], j, kBgGroup)};
            const float wj[4] = {__shfl(wxy[0], j, kBgGroup), __shfl(wxy[1], j, kBgGroup),
                                 __shfl(wxy[2], j, kBgGroup), __shfl(wxy[3], j, kBgGroup)};
            const int rj[4] = {__shfl(rows4[0], j, kBgGroup), __shfl(rows4[1], j, kBgGroup),
                               __shfl(rows4[2], j, kBgGroup), __shfl(rows4[3], j, kBgGroup)};
            // Both shuffles run in every lane: a lane whose source lane is
            // switched off by a branch would read 0.
            const float wz0j = __shfl(wz0, j, kBgGroup), wz1j = __shfl(wz1, j, kBgGroup);
            const float wzj = layer ? wz1j : wz0j;
            const int row = q == 0 ? rj[0] : (q == 1 ? rj[1] : (q == 2 ? rj[2] : rj[3]));
            const float wq = q == 0 ? wj[0] : (q == 1 ? wj[1] : (q == 2 ? wj[2] : wj[3]));
            const float vc = ch == 0 ? vj[0] : (ch == 1 ? vj[1] : (ch == 2 ? vj[2] : vj[3]));
            if (row >= 0) {
                const float val = (wq * vc) * wzj;
                if (val != 0.f) atomic_add_f32(b.grad + static_cast<int64_t>(row + layer) * 4 + ch, val);
                if (ch == 0 && b.mask != nullptr) b.mask[row + layer] = 1;
            }
        }
        lt -= __shfl(k.incl, k.stop_lane, kBgGroup);
        accum -= __shfl(acc_incl, k.stop_lane, kBgGroup);
        if (k.stopped) break;
    }
}

// Total variation of an MSI (msi_tv_grad_sparse_kernel,
// loss_kernel.cu:248-333): one thread per listed cell, all its channels.
__global__ __launch_bounds__(256) void k_msi_tv(const int32_t* __restrict__ links, int reso,
                                                const float* __restrict__ msi, int64_t rows,
                                                int nl, int nc,
                                                const int32_t* __restrict__ cells, int64_t n,
                                                float scale, float scale_last,
                                                uint8_t* __restrict__ mask,
                                                float* __restrict__ grad) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    const int idx = cells[tid];
    const int64_t total = static_cast<int64_t>(2 * reso) * reso * nl;
    if (idx < 0 || idx >= total) return;
    const int z = idx % nl;
    const int tmp = idx / nl;
    const int y = tmp % reso;
    const int x = tmp / reso;
    const int nx = x == 2 * reso - 1 ? 0 : x + 1;
    const int ny = y == reso - 1 ? 0 : y + 1;
    auto link_of = [&](int xx, int yy) {
        const int l = links[xx * reso + yy];
        return (l >= 0 && l < rows) ? l : -1;
    };
    const int l00 = link_of(x, y), l01 = link_of(x, ny), l10 = link_of(nx, y);
    const float fx = static_cast<float>(2 * reso) * (1.f / 256.f);
    const float fy = static_cast<float>(reso) * (1.f / 256.f);
    const float fz = static_cast<float>(nl) * (1.f / 256.f);
    auto val = [&](int l, int zz, int c) {
        return msi[(static_cast<int64_t>(l) * nl + zz) * nc + c];
    };
    auto add = [&](int l, int zz, int c, float v, float idelta) {
        if (l < 0 || v == 0.f) return;
        const int64_t r = static_cast<int64_t>(l) * nl + zz;
        atomic_add_f32(grad + r * nc + c, v * idelta);
        if (mask != nullptr) mask[r] = 1;
    };
    for (int c = 0; c < nc; ++c) {
        const bool last = c == nc - 1;
        const float v00 = l00 >= 0 ? val(l00, z, c) : 0.f;
        const float vz = (l00 >= 0 && z + 1 < nl) ? val(l00, z + 1, c) : (last ? 0.f : v00);
        const float v01 = l01 >= 0 ? val(l01, z, c) : 0.f;
        const float v10 = l10 >= 0 ? val(l10, z, c) : 0.f;
        float dx = v10 - v00, dy = v01 - v00, dz = vz - v00;
        const float idelta = (last ? scale_last : scale) * rsqrtf(1e-9f + dx * dx + dy * dy + dz * dz);
        dx *= fx;
        dy *= fy;
        dz *= fz;
        add(l00, z, c, -(dx + dy + dz), idelta);
        if (z + 1 < nl) add(l00, z + 1, c, dz, idelta);
        add(l01, z, c, dy, idelta);
        add(l10, z, c, dx, idelta);
    }
}

int make_bg_view(const svox2_grid_t* g, float step, float bg, float stop, BgView& b) {
    SVOX2_REQUIRE(g->background_data != nullptr, "grid.background_data is NULL");
    SVOX2_REQUIRE(g->background_reso >= 1 && g->background_reso < 32768,
                  "background_reso %d out of range", g->background_reso);
    SVOX2_REQUIRE(g->background_nlayers >= 2, "background_nlayers %d must be >= 2",
                  g->background_nlayers);
    SVOX2_REQUIRE(g->background_rows >= 0 &&
                      g->background_rows * g->background_nlayers < INT_MAX,
                  "background_rows %lld x nlayers %d exceeds int32 layer indices",
                  (long long)g->background_rows, g->background_nlayers);
    SVOX2_REQUIRE(reinterpret_cast<uintptr_t>(g->background_data) % 16 == 0,
                  "background_data must be 16-byte aligned");
    SVOX2_REQUIRE(step > 0.f, "step_size must be > 0");
    const float steps = static_cast<float>(g->background_nlayers) / step;
    SVOX2_REQUIRE(steps < 1e6f, "background needs too many steps (%g)", (double)steps);
    b.links = g->background_links;
    b.data = reinterpret_cast<const float4*>(g->background_data);
    b.grad = nullptr;
    b.mask = nullptr;
    b.rows = g->background_rows;
    b.reso = g->background_reso;
    b.nlayers = g->background_nlayers;
    b.n_steps = static_cast<int>(steps) + 2;
    for (int i = 0; i < 3; ++i) {
        b.offset[i] = g->offset[i];
        b.scaling[i] = g->scaling[i];
        b.sph[i] = 2.f / static_cast<float>(g->size[i]);
    }
    b.bg = bg;
    b.stop = stop;
    return ok();
}

}  // namespace

bool grid_has_background(const svox2_grid_t* g) {
    return g != nullptr && g->background_links != nullptr && g->background_nlayers > 0;
}

int check_background(const svox2_grid_t* g, const svox2_render_options_t* opt) {
    BgView b;
    return make_bg_view(g, opt->step_size, opt->background_brightness, opt->stop_thresh, b);
}

int launch_background(const svox2_grid_t* g, const svox2_render_options_t* opt,
                      const float* origins, const float* dirs, const svox2_camera_t* cam,
                      int64_t n, const float* log_t_fg, float* rgb, float* bg_rgb_out,
                      hipStream_t s) {
    BgView b;
    int st = make_bg_view(g, opt->step_size, opt->background_brightness, opt->stop_thresh, b);
    if (st) return st;
    if (n == 0) return ok();
    const unsigned blocks = blocks_for(n * kBgGroup, kBgBlock);
    if (cam != nullptr)
        hipLaunchKernelGGL(k_background<true>, dim3(blocks), dim3(kBgBlock), 0, s, b, origins,
                           dirs, *cam, n, log_t_fg, rgb, bg_rgb_out);
    else
        hipLaunchKernelGGL(k_background<false>, dim3(blocks), dim3(kBgBlock), 0, s, b, origins,
                           dirs, svox2_camera_t{}, n, log_t_fg, rgb, bg_rgb_out);
    return check_launch("background");
}

int launch_background_backward(const svox2_grid_t* g, const svox2_render_options_t* opt,
                               const float* origins, const float* dirs, int64_t n,
                               const float* grad_in, const float* color_cache, int rgb_residual,
                               float norm, const float* log_t_fg, const float* bg_src,
                               int bg_src_is_fg, float sparsity, const svox2_grads_t* grads,
                               hipStream_t s) {
    BgView b;
    int st = make_bg_view(g, opt->step_size, opt->background_brightness, opt->stop_thresh, b);
    if (st) return st;
    if (grads == nullptr || grads->grad_background_out == nullptr || n == 0) return ok();
    b.grad = grads->grad_background_out;
    b.mask = grads->mask_background_out;
    const unsigned blocks = blocks_for(n * kBgGroup, kBgBlock);
    hipLaunchKernelGGL(k_background_backward, dim3(blocks), dim3(kBgBlock), 0, s, b, origins,
                       dirs, n, grad_in, color_cache, rgb_residual, norm, log_t_fg, bg_src,
                       bg_src_is_fg, sparsity);
    return check_launch("background backward");
}

}  // namespace svox2amd

using namespace svox2amd;

extern "C" int svox2_msi_tv_grad_sparse(const int32_t* links, int32_t reso, const float* msi,
                                        int64_t msi_rows, int32_t nlayers, int32_t nchannels,
                                        const int32_t* rand_cells, int64_t n_cells,
                                        uint8_t* mask_out, float scale, float scale_last,
                                        float* grad_msi, void* stream) {
    SVOX2_REQUIRE(reso >= 1 && reso < 32768, "reso %d out of range", reso);
    SVOX2_REQUIRE(nlayers >= 1 && nchannels >= 1, "nlayers / nchannels must be >= 1");
    SVOX2_REQUIRE(msi_rows >= 0 && msi_rows < INT_MAX, "msi_rows out of range");
    SVOX2_REQUIRE(static_cast<int64_t>(2 * reso) * reso * nlayers < INT_MAX,
                  "MSI has too many cells for int32 indices");
    SVOX2_REQUIRE(n_cells >= 0, "negative cell count");
    if (n_cells == 0) return ok();
    SVOX2_REQUIRE(links != nullptr && rand_cells != nullptr && grad_msi != nullptr &&
                      (msi != nullptr || msi_rows == 0),
                  "links / msi / rand_cells / grad_msi are NULL");
    const float nl = static_cast<float>(n_cells);
    hipLaunchKernelGGL(k_msi_tv, dim3(blocks_for(n_cells, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), links, reso, msi, msi_rows, nlayers,
                       nchannels, rand_cells, n_cells, scale / nl, scale_last / nl, mask_out,
                       grad_msi);
    return check_launch("msi_tv_grad_sparse");
}
