/*
 * TEST INFRASTRUCTURE ONLY - the oracle.  Nothing in the product (svox2_amd/)
 * may link or call this file; only tests/, the smoke check and the CPU
 * baseline leg of bench.py use it, as the checker.
 *
 * A plain C, single-threaded, FP32 restatement of the reference's trilinear
 * ray march (svox2/csrc/render_lerp_kernel_cuvol.cu:30-120, trace_ray_cuvol)
 * and its per-ray set-up (svox2/csrc/include/render_util.cuh:578-628,
 * ray_find_bounds; :234-295, compute_skip_dist; :71-91, trilerp_cuvol_one;
 * include/cuda_util.cuh:61-75, transform_coord / lerp).  Built with
 * -ffp-contract=off and explicit fmaf so every index-determining value (grid
 * origin and direction, tmin/tmax, sample position, cell, fractional offset,
 * skip distance, trilinear density, log transmittance) rounds exactly as the
 * GPU kernel's.  Two deliberate, documented choices shared with the kernel:
 *   - 1/|d| is computed as 1.0f / sqrtf(d.d) (the reference uses CUDA's
 *     rnorm3df, whose last-bit behaviour is CUDA's own);
 *   - the stop test exp(logT) < stop_thresh is evaluated exactly as
 *     logT < L, L = log(stop_thresh) rounded up to float.
 *
 * Entry points:
 *   ref_render   per-ray colour (FP32), final log transmittance and the
 *                (iterations, evaluated, contributing, FNV-1a hash) counters
 *                the GPU stats kernel produces;
 *   ref_samples  every evaluated (not skipped) sample of every ray, ignoring
 *                the stop test, for the FP64 oracle to integrate;
 *   ref_svox1_samples  the svox1 backend's nearest-neighbour walk, every
 *                occupied-cell sample with its step and stop flags;
 *   ref_grid_weight_render  per-voxel maximum weight of a camera's rays
 *                through a dense density volume (resample's weight render);
 *   ref_fused    forward colour plus the MSE gradient (render_ray_backward_kernel
 *                with grad_out_is_rgb, render_lerp_kernel_cuvol.cu:692-773, and
 *                trace_ray_cuvol_backward :227-383), scattering every
 *                contributing sample into the 8 corner rows as the reference
 *                does (render_util.cuh:93-153): the scalar CPU port timed as
 *                the CPU baseline (rays split over n_threads OpenMP threads, the
 *                gradient updates atomic; one thread is the sequential order).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct {
    const int32_t* links;  /* size0 x size1 x size2, row-major */
    const float* density;  /* rows */
    const float* sh;       /* rows x 3*basis_dim, may be NULL for ref_samples */
    int64_t rows;
    int32_t size[3];
    int32_t basis_dim;
    float offset[3];
    float scaling[3];
} ref_grid_t;

typedef struct {
    float bg;
    float step;
    float sigma_thresh;
    float stop_thresh;
    float near_clip;
    int32_t spheric;
    int32_t last_opaque;
    int32_t skip; /* 1: empty-space skipping as the kernel; 0: never skip */
} ref_opt_t;

typedef struct {
    float o[3], d[3], inv[3];
    float ws, tmin, tmax;
} ray_t;

static float log_stop_threshold(float s) {
    if (!(s > 0.f)) return -INFINITY;
    double L = log((double)s);
    float f = (float)L;
    if ((double)f < L) f = nextafterf(f, INFINITY);
    return f;
}

static float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void setup_ray(const ref_grid_t* g, const ref_opt_t* o, const float* wo, const float* wd,
                      const float* clip, ray_t* r) {
    float d[3];
    for (int i = 0; i < 3; ++i) {
        r->o[i] = fmaf(wo[i], g->scaling[i], g->offset[i]);
        d[i] = wd[i] * g->scaling[i];
    }
    const float delta_scale = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int i = 0; i < 3; ++i) {
        r->d[i] = d[i] * delta_scale;
        r->inv[i] = (float)(1.0 / (double)r->d[i]);
    }
    r->ws = delta_scale * o->step;
    if (o->spheric) {
        float so[3], sd[3];
        for (int i = 0; i < 3; ++i) {
            const float s = 2.f / (float)g->size[i];
            so[i] = fmaf(r->o[i] + 0.5f, s, -1.f);
            sd[i] = r->d[i] * s;
        }
        const float q2a = 2.f * dot3(sd, sd);
        const float qb = 2.f * dot3(so, sd);
        const float f = qb * qb - 2.f * q2a * dot3(so, so);
        const float det_far = f + 2.f * q2a * 1.f * 1.f;
        const float rn = 1.f - o->near_clip;
        const float det_near = f + 2.f * q2a * rn * rn;
        if (det_far < 0.f || det_near < 0.f) {
            r->tmin = 1e-9f;
            r->tmax = 0.f;
        } else {
            r->tmax = (-qb + sqrtf(det_far)) / q2a;
            r->tmin = (-qb - sqrtf(det_near)) / q2a;
        }
    } else {
        r->tmin = o->near_clip / r->ws * o->step;
        r->tmax = 2e3f;
        for (int i = 0; i < 3; ++i) {
            const float t1 = (-0.5f - r->o[i]) * r->inv[i];
            const float t2 = ((float)g->size[i] - 0.5f - r->o[i]) * r->inv[i];
            if (r->d[i] != 0.f) {
                r->tmin = fmaxf(r->tmin, fminf(t1, t2));
                r->tmax = fminf(r->tmax, fmaxf(t1, t2));
            }
        }
    }
    if (clip) {
        r->tmin = fmaxf(r->tmin, clip[0] / r->ws * o->step);
        r->tmax = fminf(r->tmax, clip[1] / r->ws * o->step);
    }
}

static float skip_dist(int link, const int* l, const float* p, const ray_t* r) {
    if (link >= -1) return 0.f;
    const uint32_t shift = (uint32_t)(-(int64_t)link) - 1u;
    const uint32_t side = (uint32_t)((float)(1 << shift) - 1.f);
    float tmin = 0.f, tmax = 1e9f;
    for (int i = 0; i < 3; ++i) {
        const int ul = ((l[i] >> shift) << shift) - l[i];
        const float t1 = ((float)ul - p[i]) * r->inv[i];
        const float t2 = ((float)((uint32_t)ul + side) - p[i]) * r->inv[i];
        if (r->d[i] != 0.f) {
            tmin = fmaxf(tmin, fminf(t1, t2));
            tmax = fminf(tmax, fmaxf(t1, t2));
        }
    }
    return (tmin > 0.f) ? 0.f : tmax;
}

static float lerpf(float a, float b, float w) { return fmaf(w, b - a, a); }

static float trilerp8(const float* v, const float* p) {
    const float ix0y0 = lerpf(v[0], v[1], p[2]);
    const float ix0y1 = lerpf(v[2], v[3], p[2]);
    const float ix0 = lerpf(ix0y0, ix0y1, p[1]);
    const float ix1y0 = lerpf(v[4], v[5], p[2]);
    const float ix1y1 = lerpf(v[6], v[7], p[2]);
    const float ix1 = lerpf(ix1y0, ix1y1, p[1]);
    return lerpf(ix0, ix1, p[0]);
}

static int sample_cell(const ref_grid_t* g, const ray_t* r, float t, float* p, int* l) {
    for (int i = 0; i < 3; ++i) {
        float x = fmaf(t, r->d[i], r->o[i]);
        x = fminf(fmaxf(x, 0.f), (float)g->size[i] - 1.f);
        int li = (int)x;
        if (li > g->size[i] - 2) li = g->size[i] - 2;
        l[i] = li;
        p[i] = x - (float)li;
    }
    return l[0] * g->size[1] * g->size[2] + l[1] * g->size[2] + l[2];
}

static void corner_links(const ref_grid_t* g, int base, int* lk) {
    const int sx = g->size[1] * g->size[2], sy = g->size[2];
    const int off[8] = {0, 1, sy, sy + 1, sx, sx + 1, sx + sy, sx + sy + 1};
    for (int k = 0; k < 8; ++k) lk[k] = g->links[base + off[k]];
}

static int valid_link(const ref_grid_t* g, int lk) { return lk >= 0 && (int64_t)lk < g->rows; }

static float sh_basis(int j, float x, float y, float z) {
    switch (j) {
        case 0: return 0.28209479177387814f;
        case 1: return -0.4886025119029199f * y;
        case 2: return 0.4886025119029199f * z;
        case 3: return -0.4886025119029199f * x;
        case 4: return 1.0925484305920792f * (x * y);
        case 5: return -1.0925484305920792f * (y * z);
        case 6: return 0.31539156525252005f * (2.f * (z * z) - (x * x) - (y * y));
        case 7: return -1.0925484305920792f * (x * z);
        default: return 0.5462742152960396f * ((x * x) - (y * y));
    }
}

#define FNV_MIX(h, w) ((h) = ((h) ^ (uint32_t)(w)) * 16777619u)

static uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* One ray: colour (FP32), final log transmittance, counters. */
static void render_one(const ref_grid_t* g, const ref_opt_t* o, float log_stop, const float* wo,
                       const float* wd, const float* clip, float* rgb, float* log_t_out,
                       int32_t* stats) {
    const int B = g->basis_dim, C = 3 * B;
    float basis[9];
    for (int j = 0; j < B; ++j) basis[j] = sh_basis(j, wd[0], wd[1], wd[2]);
    ray_t r;
    setup_ray(g, o, wo, wd, clip, &r);
    float out[3] = {0.f, 0.f, 0.f};
    float log_t = 0.f;
    uint32_t n_iter = 0, n_eval = 0, n_hit = 0, h = 2166136261u;
    if (r.tmin <= r.tmax) {
        float t = r.tmin, ws = r.ws;
        int guard = 0;
        while (t <= r.tmax && ++guard < (1 << 22)) {
            ++n_iter;
            float p[3];
            int l[3], lk[8];
            const int base = sample_cell(g, &r, t, p, l);
            const int link0 = g->links[base];
            if (o->skip && link0 < -1) {
                const float skip = skip_dist(link0, l, p, &r);
                if (skip >= o->step) {
                    t += ceilf(skip / o->step) * o->step;
                    continue;
                }
            }
            ++n_eval;
            FNV_MIX(h, base);
            FNV_MIX(h, f2u(t));
            corner_links(g, base, lk);
            float dn[8];
            for (int k = 0; k < 8; ++k) dn[k] = valid_link(g, lk[k]) ? g->density[lk[k]] : 0.f;
            const float sigma = trilerp8(dn, p);
            if (o->last_opaque && t + o->step > r.tmax) ws = 1e9f;
            if (sigma > o->sigma_thresh) {
                ++n_hit;
                FNV_MIX(h, 1u);
                if (g->sh != NULL && rgb != NULL) {
                    float col[3] = {0.f, 0.f, 0.f};
                    for (int ch = 0; ch < C; ++ch) {
                        float v[8];
                        for (int k = 0; k < 8; ++k)
                            v[k] = valid_link(g, lk[k]) ? g->sh[(int64_t)lk[k] * C + ch] : 0.f;
                        col[ch / B] += trilerp8(v, p) * basis[ch % B];
                    }
                    const float pcnt = ws * sigma;
                    const float weight = expf(log_t) * (1.f - expf(-pcnt));
                    log_t -= pcnt;
                    for (int c = 0; c < 3; ++c) out[c] += weight * fmaxf(col[c] + 0.5f, 0.f);
                } else {
                    log_t -= ws * sigma;
                }
                if (log_t < log_stop) {
                    log_t = -1e3f;
                    break;
                }
            } else {
                FNV_MIX(h, 0u);
            }
            t += o->step;
        }
    }
    const float tb = expf(log_t) * o->bg;
    for (int c = 0; c < 3; ++c) out[c] += tb;
    FNV_MIX(h, f2u(log_t));
    if (rgb) memcpy(rgb, out, sizeof(out));
    if (log_t_out) *log_t_out = log_t;
    if (stats) {
        stats[0] = (int32_t)n_iter;
        stats[1] = (int32_t)n_eval;
        stats[2] = (int32_t)n_hit;
        stats[3] = (int32_t)h;
    }
}

int ref_render(const ref_grid_t* g, const float* origins, const float* dirs, const float* clip,
               int64_t n, const ref_opt_t* o, float* rgb_out, float* log_t_out,
               int32_t* stats_out) {
    const float log_stop = log_stop_threshold(o->stop_thresh);
    for (int64_t i = 0; i < n; ++i) {
        render_one(g, o, log_stop, origins + 3 * i, dirs + 3 * i, clip ? clip + 2 * i : NULL,
                   rgb_out ? rgb_out + 3 * i : NULL, log_t_out ? log_t_out + i : NULL,
                   stats_out ? stats_out + 4 * i : NULL);
    }
    return 0;
}

/* Every evaluated sample of rays [0, n) in ray order, up to `cap` entries;
 * returns the total count (callers re-run with a larger cap if it exceeds
 * cap).  Per sample: ray id, t, lower cell l[3], fractional offset p[3], the
 * world step used for its opacity (1e9 for a last-opaque sample) and flags:
 * bit 0 = the kernel's FP32 density exceeds sigma_thresh (the sample
 * contributes), bit 1 = the kernel's FP32 march has not stopped before it.
 * Samples after the stop are still emitted (flag bit 1 clear) so that an FP64
 * integrator can make its own stop decision. */
int64_t ref_samples(const ref_grid_t* g, const float* origins, const float* dirs,
                    const float* clip, int64_t n, const ref_opt_t* o, int64_t cap,
                    int32_t* ray_out, float* t_out, int32_t* l_out, float* p_out,
                    float* ws_out, uint8_t* flag_out) {
    const float log_stop = log_stop_threshold(o->stop_thresh);
    int64_t count = 0;
    for (int64_t i = 0; i < n; ++i) {
        ray_t r;
        setup_ray(g, o, origins + 3 * i, dirs + 3 * i, clip ? clip + 2 * i : NULL, &r);
        if (r.tmin > r.tmax) continue;
        float t = r.tmin, ws = r.ws, log_t = 0.f;
        int stopped = 0;
        int guard = 0;
        while (t <= r.tmax && ++guard < (1 << 22)) {
            float p[3];
            int l[3], lk[8];
            const int base = sample_cell(g, &r, t, p, l);
            const int link0 = g->links[base];
            if (o->skip && link0 < -1) {
                const float skip = skip_dist(link0, l, p, &r);
                if (skip >= o->step) {
                    t += ceilf(skip / o->step) * o->step;
                    continue;
                }
            }
            corner_links(g, base, lk);
            float dn[8];
            for (int k = 0; k < 8; ++k) dn[k] = valid_link(g, lk[k]) ? g->density[lk[k]] : 0.f;
            const float sigma = trilerp8(dn, p);
            if (o->last_opaque && t + o->step > r.tmax) ws = 1e9f;
            const int contrib = sigma > o->sigma_thresh;
            if (count < cap) {
                ray_out[count] = (int32_t)i;
                t_out[count] = t;
                for (int a = 0; a < 3; ++a) {
                    l_out[3 * count + a] = l[a];
                    p_out[3 * count + a] = p[a];
                }
                ws_out[count] = ws;
                flag_out[count] = (uint8_t)(contrib | (stopped ? 0 : 2));
            }
            ++count;
            if (contrib && !stopped) {
                log_t -= ws * sigma;
                if (log_t < log_stop) stopped = 1;
            }
            t += o->step;
        }
    }
    return count;
}

/* The svox1 backend's nearest-neighbour walk (render_svox1_kernel.cu:50-160,
 * trace_ray; compute_skip_dist_nn :23-48; _intersect_aabb_unit,
 * render_util.cuh:448-461), in FP32 as the kernel: the grid origin shifted by
 * +0.5, t from max(0, box entry), each occupied cell a sample whose step is
 * the distance to the cell's exit plus 1e-2, each empty block (link < 0,
 * 2^(-link - 1) cells on a side) skipped to its exit plus 1e-2.  Every
 * occupied-cell sample up to the forward's stop is listed: ray, link, world
 * step dt (step times the world scale) and flag bits: 1 density (+1e9 for an
 * opaque last sample) above sigma_thresh; 2 the forward stops after it (its
 * transmittance before it <= stop_thresh); 4 the backward stops after it
 * (transmittance after it <= stop_thresh); 8 opaque last sample.  Links >=
 * rows count as empty single cells.  Returns the sample count (entries past
 * cap are not written). */
int64_t ref_svox1_samples(const ref_grid_t* g, const float* origins, const float* dirs,
                          int64_t n, const ref_opt_t* o, int64_t cap, int32_t* ray_out,
                          int32_t* link_out, float* dt_out, uint8_t* flag_out) {
    int64_t count = 0;
    for (int64_t i = 0; i < n; ++i) {
        float og[3], d[3], inv[3];
        for (int a = 0; a < 3; ++a) {
            og[a] = fmaf(origins[3 * i + a], g->scaling[a], g->offset[a]);
            d[a] = dirs[3 * i + a] * g->scaling[a];
        }
        const float ds = 1.0f / sqrtf(dot3(d, d));
        for (int a = 0; a < 3; ++a) {
            d[a] *= ds;
            inv[a] = d[a] == 0.f ? 1e9f : (float)(1.0 / (double)d[a]);
        }
        float t = 0.f, tmax = 1e9f;
        for (int a = 0; a < 3; ++a) {
            og[a] += 0.5f;
            const float t1 = (0.0f - og[a]) * inv[a];
            const float t2 = ((float)g->size[a] - 1.f - og[a]) * inv[a];
            t = fmaxf(t, fminf(t1, t2));
            tmax = fminf(tmax, fmaxf(t1, t2));
        }
        if (t > tmax) continue;
        float log_t = 0.f;
        int guard = 0;
        while (t < tmax && ++guard < (1 << 22)) {
            float p[3];
            int l[3];
            for (int a = 0; a < 3; ++a) {
                p[a] = og[a] + t * d[a];
                p[a] = fminf(fmaxf(p[a], 0.f), (float)g->size[a] - 1.f);
                l[a] = (int)p[a];
                if (l[a] > g->size[a] - 1) l[a] = g->size[a] - 1;
                p[a] -= (float)l[a];
            }
            const int link = g->links[((int64_t)l[0] * g->size[1] + l[1]) * g->size[2] + l[2]];
            if (link >= 0 && (int64_t)link < g->rows) {
                float dt = 1e9f;
                for (int a = 0; a < 3; ++a) {
                    const float t1 = -p[a] * inv[a];
                    const float t2 = t1 + inv[a];
                    dt = fminf(dt, fmaxf(t1, t2));
                }
                dt += 1e-2f;
                t += dt;
                float sigma = g->density[link];
                int flag = 0;
                if (o->last_opaque && t + o->step > tmax) {
                    sigma += 1e9f;
                    flag |= 8;
                }
                if (sigma > o->sigma_thresh) {
                    flag |= 1;
                    const float lt = -(dt * ds) * sigma;
                    const float tr = expf(log_t);
                    log_t += lt;
                    if (tr <= o->stop_thresh) flag |= 2;
                    if (expf(log_t) <= o->stop_thresh) flag |= 4;
                }
                if (count < cap) {
                    ray_out[count] = (int32_t)i;
                    link_out[count] = link;
                    dt_out[count] = dt * ds;
                    flag_out[count] = (uint8_t)flag;
                }
                ++count;
                if (flag & 2) break;
            } else {
                const uint32_t dist = link < 0 ? (uint32_t)(-link) : 1u;
                const uint32_t shift = dist - 1 > 30 ? 30 : dist - 1;
                const uint32_t side = 1u << shift;
                float smax = 1e9f;
                for (int a = 0; a < 3; ++a) {
                    const int ul = ((l[a] >> shift) << shift) - l[a];
                    const float t1 = ((float)ul - p[a]) * inv[a];
                    const float t2 = ((float)(int64_t)(ul + (int64_t)side) - p[a]) * inv[a];
                    if (d[a] != 0.f) smax = fminf(smax, fmaxf(t1, t2));
                }
                t += fmaxf(smax, 0.f) + 1e-2f;
            }
        }
    }
    return count;
}

/* Per-ray geometry: grid origin, unit grid direction, world step, tmin, tmax
 * (9 floats per ray) exactly as the kernel computes them. */
int ref_ray_setup(const ref_grid_t* g, const float* origins, const float* dirs,
                  const float* clip, int64_t n, const ref_opt_t* o, float* out9) {
    for (int64_t i = 0; i < n; ++i) {
        ray_t r;
        setup_ray(g, o, origins + 3 * i, dirs + 3 * i, clip ? clip + 2 * i : NULL, &r);
        float* q = out9 + 9 * i;
        q[0] = r.o[0]; q[1] = r.o[1]; q[2] = r.o[2];
        q[3] = r.d[0]; q[4] = r.d[1]; q[5] = r.d[2];
        q[6] = r.ws; q[7] = r.tmin; q[8] = r.tmax;
    }
    return 0;
}

/* World ray of pixel (ix, iy) of a pinhole camera (c2w 3x4 row-major), as
 * cam2world_ray (render_util.cuh:526-544) without the NDC warp. */
static void pixel_ray(const float* c2w, float fx, float fy, float cx, float cy, int ix, int iy,
                      float* origin, float* dir) {
    float x = ((float)ix + 0.5f - cx) / fx;
    float y = ((float)iy + 0.5f - cy) / fy;
    float z = sqrtf((float)((double)(x * x + y * y) + 1.0));
    x /= z;
    y /= z;
    z = 1.0f / z;
    dir[0] = c2w[0] * x + c2w[1] * y + c2w[2] * z;
    dir[1] = c2w[4] * x + c2w[5] * y + c2w[6] * z;
    dir[2] = c2w[8] * x + c2w[9] * y + c2w[10] * z;
    origin[0] = c2w[3];
    origin[1] = c2w[7];
    origin[2] = c2w[11];
}

/* All pixel rays of a camera, row-major. */
int ref_camera_rays(const float* c2w, float fx, float fy, float cx, float cy, int32_t width,
                    int32_t height, float* origins, float* dirs) {
    for (int iy = 0; iy < height; ++iy) {
        for (int ix = 0; ix < width; ++ix) {
            const int64_t i = (int64_t)iy * width + ix;
            pixel_ray(c2w, fx, fy, cx, cy, ix, iy, origins + 3 * i, dirs + 3 * i);
        }
    }
    return 0;
}

/* Per-voxel maximum compositing weight of one camera's pixel rays through
 * the dense float volume data (size[0] x size[1] x size[2]), raised into out
 * (same shape, non-negative): grid_trace_ray / grid_weight_render_kernel
 * (misc_kernel.cu:184-334,418-445).  Sequential per ray: t from the volume
 * entry in steps of step_size (no skip), position fmaf(t, d, o) clamped,
 * trilinear density; a sample with density > 1e-8 has weight
 * exp(logT) (1 - exp(-ws sigma)), raises the 8 corners of its cell, and
 * stops the ray when logT falls below log(stop_thresh).  With
 * last_sample_opaque the last sample gets density + 1e9 and logT reset to 0. */
int ref_grid_weight_render(const float* data, const int32_t* size, const float* c2w, float fx,
                           float fy, float cx, float cy, int32_t width, int32_t height,
                           float step_size, float stop_thresh, int32_t last_opaque,
                           const float* offset, const float* scaling, float* out) {
    ref_grid_t g;
    memset(&g, 0, sizeof(g));
    for (int i = 0; i < 3; ++i) {
        g.size[i] = size[i];
        g.offset[i] = offset[i];
        g.scaling[i] = scaling[i];
    }
    ref_opt_t o;
    memset(&o, 0, sizeof(o));
    o.step = step_size;
    const float log_stop = log_stop_threshold(stop_thresh);
    const int64_t sx = (int64_t)size[1] * size[2], sy = size[2];
    for (int iy = 0; iy < height; ++iy) {
        for (int ix = 0; ix < width; ++ix) {
            float wo[3], wd[3];
            pixel_ray(c2w, fx, fy, cx, cy, ix, iy, wo, wd);
            ray_t r;
            setup_ray(&g, &o, wo, wd, NULL, &r);
            if (r.tmin > r.tmax) continue;
            float t = r.tmin, log_t = 0.f;
            int guard = 0;
            while (t <= r.tmax && ++guard < (1 << 22)) {
                float p[3];
                int l[3];
                for (int a = 0; a < 3; ++a) {
                    float x = fmaf(t, r.d[a], r.o[a]);
                    x = fminf(fmaxf(x, 0.f), (float)size[a] - 1.f);
                    l[a] = (int)x;
                    if (l[a] > size[a] - 2) l[a] = size[a] - 2;
                    p[a] = x - (float)l[a];
                }
                const int64_t base = l[0] * sx + l[1] * sy + l[2];
                const float v[8] = {data[base], data[base + 1], data[base + sy],
                                    data[base + sy + 1], data[base + sx], data[base + sx + 1],
                                    data[base + sx + sy], data[base + sx + sy + 1]};
                float sigma = trilerp8(v, p);
                if (last_opaque && t + step_size > r.tmax) {
                    sigma += 1e9f;
                    log_t = 0.f;
                }
                if (sigma > 1e-8f) {
                    const float pcnt = r.ws * sigma;
                    const float w = expf(log_t) * (1.f - expf(-pcnt));
                    log_t -= pcnt;
                    const int64_t c[8] = {base, base + 1, base + sy, base + sy + 1, base + sx,
                                          base + sx + 1, base + sx + sy, base + sx + sy + 1};
                    for (int k = 0; k < 8; ++k)
                        if (w > out[c[k]]) out[c[k]] = w;
                    if (log_t < log_stop) break;
                }
                t += step_size;
            }
        }
    }
    return 0;
}

/* Fused forward + MSE backward of rays [0, n): rgb_out (n,3), gradients
 * accumulated into grad_density (rows) / grad_sh (rows x C), mask (rows). */
int ref_fused(const ref_grid_t* g, const float* origins, const float* dirs, const float* clip,
              int64_t n, const ref_opt_t* o, const float* rgb_gt, float beta_loss,
              float sparsity_loss, float* rgb_out, float* grad_density, float* grad_sh,
              uint8_t* mask, int32_t n_threads) {
    const float log_stop = log_stop_threshold(o->stop_thresh);
    const int B = g->basis_dim, C = 3 * B;
    const float norm = 2.f / (float)(3 * n);
    const float beta = beta_loss / (float)n;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads) if (n_threads > 1)
    for (int64_t i = 0; i < n; ++i) {
        const float* wo = origins + 3 * i;
        const float* wd = dirs + 3 * i;
        const float* cl = clip ? clip + 2 * i : NULL;
        float rgb[3], log_t_fwd;
        render_one(g, o, log_stop, wo, wd, cl, rgb, &log_t_fwd, NULL);
        memcpy(rgb_out + 3 * i, rgb, sizeof(rgb));
        float gq[3];
        for (int c = 0; c < 3; ++c) gq[c] = (rgb[c] - rgb_gt[3 * i + c]) * norm;
        float accum = fmaf(rgb[0], gq[0], fmaf(rgb[1], gq[1], rgb[2] * gq[2]));
        ray_t r;
        setup_ray(g, o, wo, wd, cl, &r);
        if (r.tmin > r.tmax) continue;
        if (beta > 0.f) {
            const float tin = expf(log_t_fwd);
            accum += beta * (1.f - tin / (1.f - tin + 1e-3f));
        }
        float basis[9];
        for (int j = 0; j < B; ++j) basis[j] = sh_basis(j, wd[0], wd[1], wd[2]);
        float t = r.tmin, ws = r.ws, log_t = 0.f;
        int guard = 0;
        while (t <= r.tmax && ++guard < (1 << 22)) {
            float p[3];
            int l[3], lk[8];
            const int base = sample_cell(g, &r, t, p, l);
            const int link0 = g->links[base];
            if (o->skip && link0 < -1) {
                const float skip = skip_dist(link0, l, p, &r);
                if (skip >= o->step) {
                    t += ceilf(skip / o->step) * o->step;
                    continue;
                }
            }
            corner_links(g, base, lk);
            float dn[8];
            for (int k = 0; k < 8; ++k) dn[k] = valid_link(g, lk[k]) ? g->density[lk[k]] : 0.f;
            const float sigma = trilerp8(dn, p);
            if (o->last_opaque && t + o->step > r.tmax) ws = 1e9f;
            if (sigma > o->sigma_thresh) {
                float col[3] = {0.f, 0.f, 0.f};
                for (int ch = 0; ch < C; ++ch) {
                    float v[8];
                    for (int k = 0; k < 8; ++k)
                        v[k] = valid_link(g, lk[k]) ? g->sh[(int64_t)lk[k] * C + ch] : 0.f;
                    col[ch / B] += trilerp8(v, p) * basis[ch % B];
                }
                const float pcnt = ws * sigma;
                const float weight = expf(log_t) * (1.f - expf(-pcnt));
                log_t -= pcnt;
                float total = 0.f, gc[3];
                for (int c = 0; c < 3; ++c) {
                    const float x = col[c] + 0.5f;
                    total += fmaxf(x, 0.f) * gq[c];
                    gc[c] = weight * (x >= 0.f ? 1.f : 0.f) * gq[c];
                }
                accum -= weight * total;
                float gs = ws * (total * expf(log_t) - accum);
                if (sparsity_loss > 0.f)
                    gs += sparsity_loss * (4.f * sigma / (1.f + 2.f * (sigma * sigma)));
                const float ax = 1.f - p[0], ay = 1.f - p[1], az = 1.f - p[2];
                for (int k = 0; k < 8; ++k) {
                    if (!valid_link(g, lk[k])) continue;
                    const float wk = ((k & 2) ? p[1] : ay) * ((k & 1) ? p[2] : az) *
                                     ((k & 4) ? p[0] : ax);
                    if (grad_density) {
#pragma omp atomic
                        grad_density[lk[k]] += wk * gs;
                    }
                    if (mask) mask[lk[k]] = 1;
                    if (grad_sh)
                        for (int ch = 0; ch < C; ++ch) {
                            const float v = wk * basis[ch % B] * gc[ch / B];
#pragma omp atomic
                            grad_sh[(int64_t)lk[k] * C + ch] += v;
                        }
                }
                if (log_t < log_stop) break;
            }
            t += o->step;
        }
    }
    return 0;
}
