// Trilinear sparse-voxel ray march (forward, backward, fused MSE, image, stats)
// for gfx950.
//
// Work mapping: one 32-lane half of a 64-wide wavefront per ray, two rays per
// wavefront.  The march is sample-parallel inside the ray: a chunk puts up to
// 32 consecutive samples on the 32 lanes, so the corner links and densities of
// all of them are fetched by two waves of independent loads instead of one
// dependent round trip per sample, and the transmittance chain is replayed in
// registers.  Colours of the live samples are then gathered four samples at a
// time with lane = (sample, corner): each lane reads one contiguous SH row and
// projects it on the SH basis; the 8 corners are summed with lane shuffles.
//
// The backward pass rematerialises the same chunks, forms every live sample's
// gradient scalars sample-parallel (the reference's running `accum` becomes a
// prefix sum over the chunk), and adds them in sample order into registers
// with lane = (corner, R/G/B/density): a ray that stays in a cell, or moves to
// a neighbour (the shared corners are kept, a lane permutation), accumulates
// without memory traffic; a corner it leaves is flushed with one no-return
// atomic instruction covering the corner's 27 SH channels and its density
// (one 128-byte row in the interleaved layout).
//
// Launches use 64-thread workgroups (one wavefront), so the hardware refills a
// CU as soon as any wavefront's two rays end.  The fused call records each
// ray's chunk count in the forward and, for large batches, runs the backward
// over the rays in descending order of it.
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
#include <string.h>
#include <mutex>
#include <type_traits>

namespace svox2amd {
namespace {

constexpr int kGroup = 32;    // lanes per ray
#ifndef SVOX2_MARCH_BLOCK
#define SVOX2_MARCH_BLOCK 64
#endif
constexpr int kBlock = SVOX2_MARCH_BLOCK;  // threads per workgroup
constexpr int kMaxIters = 1 << 22;  // hard per-ray iteration cap (termination guard)

struct OptView {
    float bg;
    float step;
    int step_ulp_max;  // largest float exponent field e with step a whole multiple of 2^(e - 150)
    float sigma_thresh;
    float log_stop;  // stop when log transmittance < log_stop
    float stop;      // the stop threshold itself (the svox1 backend compares transmittance)
    float near_clip;
    int spheric;
    int last_opaque;
};

struct Ray {
    float o[3];    // grid-space origin
    float d[3];    // unit grid-space direction
    float inv[3];  // (float)(1.0 / (double)d[i])
    float hi[3];   // grid size - 1 per axis (the position clamp)
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
        r.hi[i] = static_cast<float>(g.size[i]) - 1.f;
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

__device__ __forceinline__ float pick3(const float* a, int i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}

struct Fnv {
    uint32_t h = 2166136261u;
    __device__ __forceinline__ void mix(uint32_t w) { h = (h ^ w) * 16777619u; }
};

// 32-lane group helpers.  A ray owns one half of the wavefront; group-local
// lane ids are 0..31 and masks are the half's 32 ballot bits.
__device__ __forceinline__ uint32_t group_ballot(bool pred) {
    const unsigned long long m = __ballot(pred);
    return static_cast<uint32_t>(m >> (threadIdx.x & 32u));
}

// Orders this wavefront's LDS writes before its later LDS reads.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over each aligned group of 8 lanes, in every lane of the group, by
// data-parallel-primitive moves (VALU, no LDS round trip): lane ^ 1 and
// lane ^ 2 by quad permutes, then the other quad by the half-row mirror
// (lane 7 - i), which after the first two steps holds the other quad's sum.
// Same additions in the same order as three xor-shuffle steps 1, 2, 4.
// Every lane of the group must be active.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ float sum8(float v) {
    v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
    v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
    v += dpp<0x141>(v);  // row_half_mirror: lane 7 - i within 8
    return v;
}

// Inclusive prefix sum over each 32-lane group by DPP moves: shifts by 1, 2,
// 4 and 8 within each 16-lane row (lanes shifted in read 0), then the second
// row of each group adds the first row's total (row broadcast of lane 15).
__device__ __forceinline__ float group_scan(float v) {
    v += dpp<0x111>(v);  // row_shr:1
    v += dpp<0x112>(v);  // row_shr:2
    v += dpp<0x114>(v);  // row_shr:4
    v += dpp<0x118>(v);  // row_shr:8
    // row_bcast:15 into rows 1 and 3; rows 0 and 2 keep 0.
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xA, 0xF, false));
    return v;
}

// Value of lane ^ BIT (BIT = 1, 2 or 4): quad permutes for 1 and 2, a
// bit-mode swizzle (xor mask 4 within 32 lanes) for 4.
template <int BIT>
__device__ __forceinline__ float xor_lanes(float v) {
    if constexpr (BIT == 1) return dpp<0xB1>(v);
    else if constexpr (BIT == 2) return dpp<0x4E>(v);
    else return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x101F));
}

// Lane j of this lane's 32-lane group, j the same in every lane of the
// wavefront (both groups read their own lane j): two scalar lane reads, no
// LDS round trip, whatever lanes are active.
__device__ __forceinline__ float group_lane(float v, int j) {
    const int bits = __float_as_int(v);
    const int a = __builtin_amdgcn_readlane(bits, j);
    const int b = __builtin_amdgcn_readlane(bits, j + 32);
    return __int_as_float((threadIdx.x & 32u) ? b : a);
}

// Lane j of this lane's group, j the same across each group (it may differ
// between the two groups): lanes 0 and 32 give each group's j.  A group that
// is not active reads some lane of its own half, unused.
__device__ __forceinline__ float group_bcast(float v, int j) {
    const int ja = __builtin_amdgcn_readlane(j, 0) & 31;
    const int jb = __builtin_amdgcn_readlane(j, 32) & 31;
    const int bits = __float_as_int(v);
    const int a = __builtin_amdgcn_readlane(bits, ja);
    const int b = __builtin_amdgcn_readlane(bits, jb + 32);
    return __int_as_float((threadIdx.x & 32u) ? b : a);
}
__device__ __forceinline__ int group_bcast(int v, int j) {
    return __float_as_int(group_bcast(__int_as_float(v), j));
}

// 1 + the highest set lane of pred in either group (0 if none), the same in
// every lane; inactive lanes count as unset.
__device__ __forceinline__ int wave_last_set(bool pred) {
    const unsigned long long m = __ballot(pred);
    const uint32_t a = static_cast<uint32_t>(m), b = static_cast<uint32_t>(m >> 32);
    return max(a ? 32 - static_cast<int>(__clz(a)) : 0, b ? 32 - static_cast<int>(__clz(b)) : 0);
}

// Position (0-based) of the n-th set bit of m; m must have more than n bits set.
__device__ __forceinline__ int nth_set_bit(uint32_t m, int n) {
    int pos = 0;
#pragma unroll
    for (int w = 16; w >= 1; w >>= 1) {
        const uint32_t lo = m & ((1u << w) - 1u);
        const int c = __popc(lo);
        if (n >= c) {
            n -= c;
            m >>= w;
            pos += w;
        } else {
            m = lo;
        }
    }
    return pos;
}

// Per-ray constants every lane keeps: SH basis at the view direction.
template <int B>
struct RayCtx {
    float basis[B];
    __device__ __forceinline__ explicit RayCtx(const float* vdir) {
#pragma unroll
        for (int j = 0; j < B; ++j) basis[j] = sh_basis(j, vdir[0], vdir[1], vdir[2]);
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
        x = fminf(fmaxf(x, 0.f), r.hi[i]);
        l[i] = min(static_cast<int>(x), g.size[i] - 2);
        p[i] = x - static_cast<float>(l[i]);
    }
    return l[0] * g.stride_x + l[1] * g.stride_y + l[2];
}

// Trilinear weight of corner k (bits dx dy dz) at fractional offset p.
__device__ __forceinline__ float corner_weight(int k, const float* p) {
    const float wx = (k & 4) ? p[0] : 1.f - p[0];
    const float wy = (k & 2) ? p[1] : 1.f - p[1];
    const float wz = (k & 1) ? p[2] : 1.f - p[2];
    return wy * wz * wx;
}

// One chunk of up to 32 evaluated samples of a ray, one sample per lane, in
// march order.
//
// The chunk is filled segment by segment: a segment is a run of consecutive
// steps t, t + step, ... with t advanced by the same sequence of float
// additions the sequential march performs, and it ends at the first sample
// whose cell is skipped.  That sample is not evaluated: the skip sets the start
// of the next segment (render_lerp_kernel_cuvol.cu:54-86), which continues on
// the skipped sample's lane.  Positions, cells, skips and densities are those
// of the reference march.  The log transmittance before each sample is
// replayed in sample order with the same float subtractions, so the stop
// decision (:103-106) and every later sample agree bit for bit with the
// sequential march.  Filling a chunk across skips keeps its lanes busy where
// the ray crawls along a surface (a skip every few samples): each skip costs
// one cell-map round trip instead of a chunk of its own.
struct Chunk {
    float t;          // this lane's sample t
    float p[3];       // fractional position in the cell
    int l[3];         // lower cell corner
    int base;         // flat index of l
    int lk[8];        // corner links
    float sigma;      // trilinear density
    float ws;         // world step of this sample (1e9 for a last-opaque sample)
    float lt_before;  // log transmittance before the sample
    float lt_after;   // log transmittance after the sample
    float weight;     // compositing weight T (1 - exp(-ws sigma)) of a live sample
    bool contrib;     // evaluated and sigma > sigma_thresh
    bool live;        // contributes and lies at or before the stop sample
    uint32_t live_mask;
    int n_place;      // lanes given a sample to evaluate (0 .. n_place - 1)
    uint32_t load_mask;  // lanes whose corner links are read (evaluated, cell may hold rows)
    int n_eval;       // evaluated samples (lanes 0 .. n_eval - 1)
    int n_iter;       // march iterations consumed (evaluated + one per skip)
    bool stopped;     // the stop test fired in this chunk
    bool more;        // the march continues after this chunk
    float next_t;     // start of the next chunk
    float log_t_end;  // log transmittance after the chunk
};

// Segments (skips + 1) at most per chunk: a chunk that reaches it ends early
// and the next chunk continues after the last skip.
constexpr int kMaxSegments = 16;

// True when t0, t0 + step, ..., t0 + 31 step are all exactly representable in
// the binade of t0: then the running sum t += step never rounds.  t0 must be
// a positive normal float (exponent field e in 1..253; a sign bit, zero,
// infinity or NaN fails that), step a whole multiple of the binade's ulp
// 2^(e - 150) (e <= step_ulp_max), and the rounded end point in the same
// binade (it reaches the next power of two whenever the exact one does:
// rounding is monotone and powers of two are representable).
__device__ __forceinline__ bool sum_chain_exact(float t0, float step, int step_ulp_max) {
    const uint32_t e = __float_as_uint(t0) >> 23;
    const uint32_t e_end = __float_as_uint(fmaf(31.f, step, t0)) >> 23;
    return (e - 1u) < 253u && static_cast<int>(e) <= step_ulp_max && e_end == e;
}

// Where a chunk's placement leaves the march: whether it continues after
// the chunk and where, and the skips taken (for the iteration counters).
struct Placement {
    bool more = false;
    float next_t = 0.f;
    int n_skips = 0;   // skips inside the chunk
    int my_skips = 0;  // skips before this lane's sample
};

template <typename DT>
__device__ __forceinline__ void place_chunk(const GridView<DT>& g, const OptView& o,
                                            const Ray& r, int lane, float t0, Chunk& c,
                                            Placement& pl) {
    // Place the samples.  Each round gives the lanes from seg_lane on
    // the steps of the segment starting at seg_t, reads their cells' skip
    // levels and finds the segment's end.
    int seg_lane = 0;      // first lane of the current segment
    float seg_t = t0;      // t of its first sample
    int n_skips = 0;       // skips before the current segment
    int my_skips = 0;      // skips before this lane's sample
    bool occupied = true;  // the cell may have a corner link that addresses a row
    float ti = t0;
    int n_place;           // evaluated lanes 0 .. n_place - 1
    bool more;
    float next_t = 0.f;
    for (int seg = 1;; ++seg) {
        const bool mine = lane >= seg_lane;
        if (mine) {
            // t of the k-th step of the segment, as the sequential float
            // additions t += step give it.  When every t of the segment lies
            // in the binade of seg_t (seg_t > 0 normal) and step is a whole
            // multiple of that binade's ulp, each of those additions is exact,
            // so seg_t + k * step (one fma, exact) is the same float;
            // otherwise the chain is replayed.
            const int k = lane - seg_lane;
            if (sum_chain_exact(seg_t, o.step, o.step_ulp_max)) {
                ti = fmaf(static_cast<float>(k), o.step, seg_t);
            } else {
                ti = seg_t;
#pragma unroll
                for (int kk = 1; kk < 32; ++kk) {
                    const float nx = ti + o.step;
                    ti = (kk <= k) ? nx : ti;
                }
            }
            c.base = sample_cell(g, r, ti, c.p, c.l);
            my_skips = n_skips;
        }
        const bool in_range = mine && ti <= r.tmax;
        float skip = 0.f;
        bool is_skip = false;
        if (in_range) {
            int level;
            if (g.cmap != nullptr) {
                // Skip level and occupancy from the cell map: a cache-resident
                // read in place of the 8 link reads for cells whose corners are
                // all empty.
                const uint32_t info = g.cmap[(c.l[0] >> 1) * g.mstride_x +
                                             (c.l[1] >> 1) * g.mstride_y + (c.l[2] >> 1)];
                level = static_cast<int>(info & 0xFFu);
                if (level == 255) {  // block mixes levels: read this cell's own link
                    const int l0 = g.links[c.base];
                    level = l0 < -1 ? -l0 - 1 : 0;
                }
                occupied = ((info >> (8 + ((c.l[0] & 1) << 2) + ((c.l[1] & 1) << 1) +
                                      (c.l[2] & 1))) & 1u) != 0;
            } else {
                const int l0 = g.links[c.base];
                level = l0 < -1 ? -l0 - 1 : 0;
            }
            if (level > 0) {
                skip = skip_dist(-1 - level, c.l, c.p, r);
                is_skip = skip >= o.step;
            }
        }
        const uint32_t skip_mask = group_ballot(is_skip);
        // t grows along the segment, so its in-range lanes are a prefix.
        const int range_end = seg_lane + __popc(group_ballot(in_range));
        const int first_skip = skip_mask ? (__ffs(skip_mask) - 1) : 32;
        if (first_skip < range_end) {
            const float ts = group_bcast(ti, first_skip);
            const float sk = group_bcast(skip, first_skip);
            next_t = ts + ceilf(sk / o.step) * o.step;
            ++n_skips;
            seg_lane = first_skip;
            seg_t = next_t;
            if (seg < kMaxSegments) continue;
            n_place = seg_lane;
            more = true;
            break;
        }
        n_place = range_end;
        more = range_end == 32;
        if (more) next_t = group_lane(ti, 31) + o.step;
        break;
    }
    c.t = ti;
    c.n_place = n_place;
    c.load_mask = group_ballot(lane < n_place && (g.cmap == nullptr || occupied));
    pl.more = more;
    pl.next_t = next_t;
    pl.n_skips = n_skips;
    pl.my_skips = my_skips;
}

// Corner links of the cell at flat index base for the lanes in load_mask
// (-1 for the others).
template <typename DT>
__device__ __forceinline__ void load_corner_links(const GridView<DT>& g, int lane, int base,
                                                  uint32_t load_mask, int* lk) {
    const int sx = g.stride_x, sy = g.stride_y;
    if ((load_mask >> lane) & 1u) {
        lk[0] = g.links[base];
        lk[1] = g.links[base + 1];
        lk[2] = g.links[base + sy];
        lk[3] = g.links[base + sy + 1];
        lk[4] = g.links[base + sx];
        lk[5] = g.links[base + sx + 1];
        lk[6] = g.links[base + sx + sy];
        lk[7] = g.links[base + sx + sy + 1];
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) lk[k] = -1;
    }
}

// Evaluates a placed chunk (c.t, c.base, c.p, c.l, c.n_place, c.load_mask):
// corner links (unless links_ready: c.lk already holds them) and densities,
// transmittance, stop and weights.
template <typename DT>
__device__ __forceinline__ void eval_chunk(const GridView<DT>& g, const OptView& o,
                                           const Ray& r, int lane, float log_t,
                                           const Placement& pl, Chunk& c,
                                           bool links_ready = false) {
    const int n_place = c.n_place;
    const bool eval = lane < n_place;
    const float ti = c.t;
    // Corner links and densities of the evaluated samples.
    if (!links_ready) load_corner_links(g, lane, c.base, c.load_mask, c.lk);
    // Corners without a row read row 0 and are zeroed after: the 8 loads
    // issue back to back without a branch each (a grid with rows reads in
    // bounds; one without any has no valid corner and reads nothing).
    uint32_t db[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    if (g.rows > 0) {
        bool ok[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ok[k] = eval && g.valid(c.lk[k]);
            db[k] = ld_bits(g.density, static_cast<int64_t>(ok[k] ? c.lk[k] : 0) * g.dstride);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) db[k] = ok[k] ? db[k] : 0u;
    }
    float dn[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) dn[k] = bits_value<DT>(db[k]);
    c.sigma = trilerp8(dn, c.p);
    c.ws = (o.last_opaque && ti + o.step > r.tmax) ? 1e9f : r.ws;
    c.contrib = eval && c.sigma > o.sigma_thresh;
    const float pcnt = c.contrib ? c.ws * c.sigma : 0.f;

    // Log transmittance before this lane's sample, subtracting the earlier
    // samples' optical depths one by one in sample order.  Only lanes up to
    // the last contributing one subtract anything (x - 0 == x), so the replay
    // stops there, and chunks without a contribution (empty-space crawls)
    // skip it.
    // The replay runs to the later of the two groups' last contributions
    // (the other group subtracts zeros there), with a wave-uniform j.
    float lt = log_t;
    const int n_rep = wave_last_set(c.contrib);
    for (int j = 0; j < n_rep; ++j) {
        const float pj = group_lane(pcnt, j);
        if (j < lane) lt -= pj;
    }
    c.lt_before = lt;
    c.lt_after = lt - pcnt;
    const uint32_t stop_mask = group_ballot(c.contrib && c.lt_after < o.log_stop);
    const int first_stop = stop_mask ? (__ffs(stop_mask) - 1) : 32;
    c.stopped = first_stop < 32;
    c.live = c.contrib && lane <= first_stop;
    c.live_mask = group_ballot(c.live);
    c.weight = c.live ? __expf(c.lt_before) * (1.f - __expf(-pcnt)) : 0.f;

    if (c.stopped) {
        c.n_eval = first_stop + 1;
        c.n_iter = first_stop + 1 + group_bcast(pl.my_skips, first_stop);
        c.more = false;
        c.log_t_end = -1e3f;
        c.next_t = 0.f;
        return;
    }
    c.n_eval = n_place;
    c.n_iter = n_place + pl.n_skips;
    c.log_t_end = n_place > 0 ? group_bcast(c.lt_after, n_place - 1) : log_t;
    c.more = pl.more;
    c.next_t = pl.next_t;
}

template <typename DT>
__device__ __forceinline__ void march_chunk(const GridView<DT>& g, const OptView& o,
                                            const Ray& r, int lane, float t0, float log_t,
                                            Chunk& c) {
    Placement pl;
    place_chunk(g, o, r, lane, t0, c, pl);
    eval_chunk(g, o, r, lane, log_t, pl, c);
}


// Adds coefficient ch of a row to its channel's projection on the SH basis
// (c0, c1, c2 for the three channels; coefficient j of a channel is weighted
// by basis[j]).
template <int B>
__device__ __forceinline__ void project_channel(int ch, float x, const float* basis, float& c0,
                                                float& c1, float& c2) {
    const int c = ch / B, j = ch % B;
    if (c == 0) c0 = fmaf(basis[j], x, c0);
    else if (c == 1) c1 = fmaf(basis[j], x, c1);
    else c2 = fmaf(basis[j], x, c2);
}

// Projections of row lk on the SH basis: col[c] = sum_j basis[j] * sh[c * B +
// j], accumulated in j order.  Rows that are 16-byte aligned with a stride of
// whole 16-byte units are read with 16-byte loads (one aligned 128-byte line
// for a float32 SH-2 row of stride 32, 64 bytes in bf16), and each loaded
// vector is consumed as it arrives (bf16 channels are widened one by one
// into the products), so the row is never held as C floats.
template <int B>
__device__ __forceinline__ void project_row(const GridView<float>& g, int lk, const float* basis,
                                            float& c0, float& c1, float& c2) {
    constexpr int C = 3 * B;
    const float* row = g.sh + static_cast<int64_t>(lk) * g.sstride;
    if (g.vec_rows) {
        const float4* r4 = reinterpret_cast<const float4*>(__builtin_assume_aligned(row, 16));
#pragma unroll
        for (int i = 0; i < (C + 3) / 4; ++i) {
            const float4 v = r4[i];
            const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; ++h)
                if (4 * i + h < C) project_channel<B>(4 * i + h, e[h], basis, c0, c1, c2);
        }
    } else {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) project_channel<B>(ch, row[ch], basis, c0, c1, c2);
    }
}

template <int B>
__device__ __forceinline__ void project_row(const GridView<uint16_t>& g, int lk,
                                            const float* basis, float& c0, float& c1, float& c2) {
    constexpr int C = 3 * B;
    const uint16_t* row = g.sh + static_cast<int64_t>(lk) * g.sstride;
    if (g.vec_rows) {
        const uint4* r4 = reinterpret_cast<const uint4*>(__builtin_assume_aligned(row, 16));
#pragma unroll
        for (int i = 0; i < (C + 7) / 8; ++i) {
            const uint4 v = r4[i];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                if (8 * i + 2 * h < C)
                    project_channel<B>(8 * i + 2 * h, __uint_as_float(w[h] << 16), basis, c0, c1, c2);
                if (8 * i + 2 * h + 1 < C)
                    project_channel<B>(8 * i + 2 * h + 1, __uint_as_float(w[h] & 0xffff0000u), basis,
                                       c0, c1, c2);
            }
        }
    } else {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) project_channel<B>(ch, ld(row, ch), basis, c0, c1, c2);
    }
}

// Ray ordering for the backward launch.  A ray's backward costs close to a
// fixed amount per chunk (the dependent link, density and row round trips),
// so the forward records its chunk count and, for batches several times
// larger than the rays the device holds at once, the backward starts the rays
// with the most chunks first: the rays of a wave are then of similar length
// (no half-wave idles behind a long ray) and the longest rays no longer finish
// last on an otherwise idle device.  Smaller batches run in one round, where
// pairing the longest rays in the same waves only slows them, and keep the
// input order.  When the backward replays recorded samples instead, its
// work is per record and per recorded chunk, and that count is the key.
constexpr int kCostBins = 1024;

__device__ __forceinline__ uint32_t backward_cost(uint32_t work) {
    return work < static_cast<uint32_t>(kCostBins) ? work : static_cast<uint32_t>(kCostBins - 1);
}

// Colours of a chunk's live samples, grouped by cell.  Consecutive live
// samples in the same cell form a run; they share the 8 corner rows, so
// each row is read and projected on the SH basis once per run.  Up to four
// runs per pass, lane = (run, corner): the lane loads its corner's row of
// the run's cell and keeps the three projections, then for each sample of
// the run in turn weights them trilinearly at the sample's position and sums
// the 8 corners.  emit(pass, i, act, s, col) is called by every lane for the
// i-th sample of each pass's runs: act says whether the lane's group has a
// sample s at step i, col (identical in the group's 8 lanes) is its raw
// colour.  The per-sample colour is the same float computation as a
// per-sample gather.
template <typename DT, int B, typename Emit>
__device__ __forceinline__ void chunk_run_colors(const GridView<DT>& g, const RayCtx<B>& rc,
                                                 const Chunk& c, int lane, uint32_t head_mask,
                                                 Emit&& emit) {
    const int n_heads = __popc(head_mask);
    const int k = lane & 7;
    for (int pass = 0; pass * 4 < n_heads; ++pass) {
        const int h = pass * 4 + (lane >> 3);
        const bool has = h < n_heads;
        const int src = has ? nth_set_bit(head_mask, h) : 0;
        const uint32_t later = src < 31 ? (head_mask >> (src + 1)) << (src + 1) : 0u;
        const uint32_t below_end = later ? ((1u << (__ffs(later) - 1)) - 1u) : 0xFFFFFFFFu;
        uint32_t run = has ? (c.live_mask & below_end & (0xFFFFFFFFu << src)) : 0u;
        int lk = __shfl(c.lk[0], src, kGroup);
#pragma unroll
        for (int kk = 1; kk < 8; ++kk) {
            const int v = __shfl(c.lk[kk], src, kGroup);
            lk = (k == kk) ? v : lk;
        }
        float c0 = 0.f, c1 = 0.f, c2 = 0.f;
        if (has && g.valid(lk)) project_row<B>(g, lk, rc.basis, c0, c1, c2);
        for (int i = 0;; ++i) {
            const bool act = run != 0u;
            if (!__any(act)) break;
            const int s = act ? __ffs(run) - 1 : 0;
            run &= run - 1u;
            float p[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] = __shfl(c.p[a], s, kGroup);
            const float w = corner_weight(k, p);
            float col[3];
            col[0] = sum8(c0 * w);
            col[1] = sum8(c1 * w);
            col[2] = sum8(c2 * w);
            emit(pass, i, act, s, col);
        }
    }
}

// Run heads of a chunk: live samples whose previous live sample (in lane
// order) is not in the same cell.  Runs are formed over the live samples
// alone, so a chunk whose live samples are compacted into lanes 0 .. n - 1
// (the records) has the same runs, passes and sums.
__device__ __forceinline__ uint32_t run_heads(const Chunk& c, int lane) {
    const uint32_t below = c.live_mask & ((1u << lane) - 1u);
    const int prev = below ? 31 - static_cast<int>(__clz(below)) : 0;
    const int prev_cell = __shfl(c.base, prev, kGroup);
    return group_ballot(c.live && !(below != 0u && prev_cell == c.base));
}

// Colours of a chunk's live samples: each live lane receives its own sample's
// colour + 0.5 in cs (before clamping), and, if acc is given, lanes with
// lane % 8 == 0 add weight * max(colour + 0.5, 0) of their group's samples to
// acc (the forward's composite, summed over lanes at the end of the ray).
template <typename DT, int B, bool TO_LANES>
__device__ __forceinline__ void chunk_colours(const GridView<DT>& g, const RayCtx<B>& rc,
                                              const Chunk& c, int lane, float* cs, float* acc) {
    const uint32_t heads = run_heads(c, lane);
    const uint32_t upto = (lane == 31) ? 0xFFFFFFFFu : ((2u << lane) - 1u);
    const uint32_t my_heads = heads & upto;
    // This lane's run (head index) and its place in the run.
    const int my_run = __popc(my_heads) - 1;
    const int head_pos = my_heads ? 31 - __clz(my_heads) : 0;
    const int my_idx = __popc(c.live_mask & ((1u << lane) - 1u) & (0xFFFFFFFFu << head_pos));
    chunk_run_colors<DT, B>(
        g, rc, c, lane, heads, [&](int pass, int i, bool act, int s, const float* col) {
            if (acc != nullptr) {
                const float w = __shfl(c.weight, s, kGroup);
                if (act && (lane & 7) == 0) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) acc[a] += w * fmaxf(col[a] + 0.5f, 0.f);
                }
            }
            if (TO_LANES) {
                const int grp = my_run - 4 * pass;
                const bool mine = c.live && grp >= 0 && grp < 4 && my_idx == i;
                const int from = mine ? 8 * grp : 0;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float v = __shfl(col[a], from, kGroup);
                    cs[a] = mine ? v + 0.5f : cs[a];
                }
            }
        });
}

// Chunk log of the fused call: the forward writes the placement of each of a
// ray's first kLogChunks chunks (every lane's t, the placed-lane count and the
// link-load mask), and the backward evaluates those chunks from it instead of
// placing them again (no cell-map reads or skip distances).  Chunk ci of ray
// r is rec[(r * kLogChunks + ci) * kLogWords + w]: w = 0..31 the lanes' t,
// 32 the placed-lane count, 33 the mask.  head[r] = (logged chunks, whether
// the march continued after the last of them, the t it continued at, 0).
struct ChunkLog {
    float* __restrict__ rec;
    int4* __restrict__ head;
};
constexpr int kLogChunks = 16;
constexpr int kLogWords = 36;

// Record stream of the fused call: the forward appends, per ray, a header and
// one 96-byte record per contributing sample of each chunk; the backward
// replays them instead of re-marching.  Slot s is rec[6 s .. 6 s + 5]
// (float4).  Ray r owns slots [r * fixed, (r + 1) * fixed) and continues in
// overflow blocks of kRecBlock slots taken with one atomic add.  Headers:
// (n > 0, -) n records follow; (0, next) continue at slot next; (-1, -) end.
// state[r] = 1 when the ray's stream is complete, 2 when the overflow region
// ran out (the backward then re-marches the ray).
//
// Long rays are replayed in segments: a ray with at least kRecSplitMin
// recorded chunks lists its segments m >= 1 in seg_list (one reservation per
// ray, seg_cap entries in all) and sets split; seg[r * (kRecSegs - 1) + m -
// 1] holds the header slot of its recorded chunk m * kRecSegChunks (-1 past
// its end), where segment m's backward starts.
struct RecStream {
    float4* __restrict__ rec;
    int32_t* __restrict__ state;
    // Counters, one per partition of rays (ray % parts), so that no single
    // address takes every ray's atomic: [p] overflow slots taken of opart,
    // [kRecParts + p] segment list entries of spart, [2 kRecParts + p] chunk
    // list entries of cpart; partition p owns the p-th slice of each region.
    uint32_t* __restrict__ next_block;
    int parts;
    int64_t opart, spart, cpart;
    int32_t* __restrict__ seg;
    int32_t* __restrict__ split;        // per ray: 1 if its segments are listed, else -1
    int2* __restrict__ seg_list;        // (ray, segment m >= 1), or (-1, -1)
    int64_t seg_cap;
    int64_t fixed;                      // slots per ray in the fixed region
    int64_t overflow_base;              // first overflow slot
    int64_t overflow_slots;             // slots in the overflow region
    // Deferred colours: the recorded chunks of rays whose colours are formed
    // by k_rec_colours.
    int2* __restrict__ clist;           // (ray, header slot), or (-1, -1)
    float4* __restrict__ chunk_rgb;     // per entry: the chunk's colour
    float4* __restrict__ ray_info;      // per ray: (first entry, entries, background term)
    int64_t clist_cap;
};
constexpr int kRecBlock = 64;  // >= 32 records + header + the end/jump header
constexpr int kRecParts = 64;  // counter partitions at most
constexpr int kDeferChunks = 64;  // recorded chunks a ray can list for deferred colours
constexpr int64_t kDeferCap = 8;  // chunk list entries per ray of the batch
constexpr int kRecSegChunks = 1;  // recorded chunks per replay segment
constexpr int kRecSegs = 24;      // segments of a split ray (the last runs to its end)
constexpr int kRecSplitMin = 2 * kRecSegChunks;  // recorded chunks that make a ray split

// A chunk of n records occupies slots slot .. slot + n (the header, then n
// records), stored field-major after the header's slot: field j (0..5) of
// record r is rec_field0(rec, slot)[j * n + r], so that each field's stores
// and loads over a chunk's records are one contiguous run.
__device__ __forceinline__ float4* rec_field0(float4* rec, int64_t slot) { return rec + 6 * slot + 6; }
__device__ __forceinline__ const float4* rec_field0(const float4* rec, int64_t slot) {
    return rec + 6 * slot + 6;
}

__device__ __forceinline__ void put_header(float4* rec, int64_t slot, int n, int64_t next) {
    rec[6 * slot] = make_float4(__int_as_float(n), __int_as_float(static_cast<int>(next)), 0.f,
                                0.f);
}

// Adds a chunk's colour to the ray's: the run groups' partial sums (lanes 0,
// 8, 16 and 24 of cacc) reduced across the group as (a0 + a8) + (a16 + a24),
// lane ^ 8 by a row rotation and lane ^ 16 by a bit-mode swizzle (no LDS
// addresses), then added to total in march order (every lane of the group
// gets the same sums).
__device__ __forceinline__ void add_chunk_colour(float* total, const float* cacc, float* sum) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = cacc[a];
        v += dpp<0x128>(v);  // row_ror:8 - lane ^ 8 within each 16-lane row
        v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));  // lane ^ 16
        sum[a] = group_lane(v, 0);
        total[a] += sum[a];
    }
}

// Forward march of one ray.  out[3] = composited colour (all lanes),
// returns the final log transmittance.  A chunk's colour sums its live
// samples' weighted colours (per run group, then across the group); the ray's
// colour is the sum of its chunks' in march order plus the background.
//
// DEFER (recording forward): the chunks' colours are left to k_rec_colours.
// The samples are recorded with zero colours, each recorded chunk's header
// slot is staged in slot_lds, and at the ray's end the chunks are listed in
// rs->clist; k_rec_colours forms the same colours and sums.  A ray whose
// records or list entries do not fit is marched again with the colours
// formed here and no records (state 2, re-marched by the backward).
template <typename DT, int B, bool STATS, bool RECORD = false, bool LOG = false,
          bool DEFER = false>
__device__ __forceinline__ float march_forward(const GridView<DT>& g, const OptView& o,
                                               const Ray& r, const RayCtx<B>& rc, int lane,
                                               float* out, int32_t* stats,
                                               uint32_t* cost = nullptr,
                                               const RecStream* rs = nullptr,
                                               int64_t ray = 0,
                                               const ChunkLog* log = nullptr,
                                               float* log_lds = nullptr,
                                               int32_t* slot_lds = nullptr) {
    bool log_more = false;   // the march continued after the last logged chunk
    float log_next_t = 0.f;  // and where
    // Record stream position of this ray: next free slot and end of block.
    int64_t cur = RECORD ? ray * rs->fixed : 0;
    int64_t end = RECORD ? cur + rs->fixed : 0;
    bool rec_ok = RECORD;
    int n_rec_chunks = 0;  // chunks whose records were written
    float total[3] = {0.f, 0.f, 0.f};  // colour composited so far
    uint32_t n_iter = 0, n_eval = 0, n_hit = 0;
    uint32_t n_chunks = 0;
    uint32_t n_live_all = 0, n_live_chunks = 0;  // replay work (recording only)
    Fnv fnv;
    float log_t = 0.f;
    bool defer = DEFER;
    for (int attempt = 0; attempt < (DEFER ? 2 : 1); ++attempt) {
        if (DEFER && attempt == 1) {  // the deferred march did not fit
            defer = false;
            total[0] = total[1] = total[2] = 0.f;
            n_chunks = n_live_all = n_live_chunks = 0;
        }
        log_t = 0.f;
        if (r.tmin <= r.tmax) {
            float t = r.tmin;
            int guard = 0;
            bool more = true;
            while (more && ++guard < kMaxIters) {
                Chunk c;
                march_chunk(g, o, r, lane, t, log_t, c);
                const int n_live = __popc(c.live_mask);
                if (LOG && n_chunks < static_cast<uint32_t>(kLogChunks)) {
                    // Staged in LDS and written out when the ray ends: global
                    // stores here would hold up the next chunk's loads (a load's
                    // wait also waits for every earlier store).
                    float* rec = log_lds + n_chunks * kLogWords;
                    rec[lane] = c.t;
                    if (lane == 0) {
                        rec[32] = __int_as_float(c.n_place);
                        rec[33] = __uint_as_float(c.load_mask);
                    }
                    log_more = c.more;
                    log_next_t = c.next_t;
                }
                ++n_chunks;
                if (RECORD && n_live > 0) {
                    n_live_all += n_live;
                    ++n_live_chunks;
                }
                if (n_live > 0) {
                    // Recorded chunk k = n_rec_chunks opens replay segment
                    // k / kRecSegChunks when k is a multiple of kRecSegChunks.
                    const bool seg_start = RECORD && rec_ok && n_rec_chunks > 0 &&
                                           n_rec_chunks % kRecSegChunks == 0 &&
                                           n_rec_chunks / kRecSegChunks < kRecSegs;
                    float cs[3] = {0.f, 0.f, 0.f};
                    if (!defer) {
                        float cacc[3] = {0.f, 0.f, 0.f}, sum[3];
                        chunk_colours<DT, B, RECORD>(g, rc, c, lane, cs, cacc);
                        add_chunk_colour(total, cacc, sum);
                    }
                    if (RECORD && rec_ok) {
                        // Header + n_live records; keep one slot for a jump/end header.
                        if (cur + n_live + 2 > end) {
                            const int part = static_cast<int>(ray % rs->parts);
                            uint32_t nb = 0;
                            if (lane == 0)
                                nb = atomicAdd(rs->next_block + part, static_cast<uint32_t>(kRecBlock));
                            nb = static_cast<uint32_t>(__shfl(static_cast<int>(nb), 0, kGroup));
                            if (static_cast<int64_t>(nb) + kRecBlock > rs->opart) {
                                rec_ok = false;
                            } else {
                                const int64_t nxt = rs->overflow_base + part * rs->opart + nb;
                                if (lane == 0) put_header(rs->rec, cur, 0, nxt);
                                cur = nxt;
                                end = nxt + kRecBlock;
                            }
                        }
                        if (DEFER && defer && n_rec_chunks >= kDeferChunks) rec_ok = false;
                        if (rec_ok) {
                            const int k = n_rec_chunks++;
                            if (seg_start && lane == 0)  // a replay segment may start here
                                rs->seg[ray * (kRecSegs - 1) + k / kRecSegChunks - 1] =
                                    static_cast<int32_t>(cur);
                            if (DEFER && defer && lane == 0) slot_lds[k] = static_cast<int32_t>(cur);
                            if (lane == 0) put_header(rs->rec, cur, n_live, 0);
                            if (c.live) {
                                const int rank = __popc(c.live_mask & ((1u << lane) - 1u));
                                float4* q = rec_field0(rs->rec, cur) + rank;
                                q[0] = make_float4(c.p[0], c.p[1], c.p[2], __int_as_float(c.base));
                                q[n_live] = make_float4(__int_as_float(c.l[0]), __int_as_float(c.l[1]),
                                                        __int_as_float(c.l[2]), c.weight);
                                q[2 * n_live] = make_float4(__int_as_float(c.lk[0]), __int_as_float(c.lk[1]),
                                                            __int_as_float(c.lk[2]), __int_as_float(c.lk[3]));
                                q[3 * n_live] = make_float4(__int_as_float(c.lk[4]), __int_as_float(c.lk[5]),
                                                            __int_as_float(c.lk[6]), __int_as_float(c.lk[7]));
                                q[4 * n_live] = make_float4(cs[0], cs[1], cs[2], c.lt_after);
                                q[5 * n_live] = make_float4(c.ws, c.sigma, 0.f, 0.f);
                            }
                            cur += n_live + 1;
                        }
                    }
                }
                if (STATS) {
                    n_iter += c.n_iter;
                    n_eval += c.n_eval;
                    n_hit += n_live;
                    for (int j = 0; j < c.n_eval; ++j) {
                        fnv.mix(static_cast<uint32_t>(__shfl(c.base, j, kGroup)));
                        fnv.mix(__float_as_uint(__shfl(c.t, j, kGroup)));
                        fnv.mix(__shfl(static_cast<int>(c.contrib), j, kGroup) ? 1u : 0u);
                    }
                }
                log_t = c.log_t_end;
                t = c.next_t;
                more = c.more;
                if (DEFER && defer && !rec_ok) break;  // march again, colours here
            }
        }
        if (!DEFER || !defer) break;
        if (!rec_ok) continue;  // the records did not fit
        if (n_rec_chunks > 0) {
            // List the recorded chunks: one reservation per ray.
            const int part = static_cast<int>(ray % rs->parts);
            int local = 0;
            if (lane == 0)
                local = static_cast<int>(atomicAdd(rs->next_block + 2 * kRecParts + part,
                                                   static_cast<uint32_t>(n_rec_chunks)));
            local = __shfl(local, 0, kGroup);
            const bool fits = static_cast<int64_t>(local) + n_rec_chunks <= rs->cpart;
            const int first = static_cast<int>(part * rs->cpart + local);
            wave_lds_fence();
            for (int k = lane; k < n_rec_chunks; k += kGroup) {
                if (static_cast<int64_t>(local) + k < rs->cpart)
                    rs->clist[static_cast<int64_t>(first) + k] =
                        fits ? make_int2(static_cast<int>(ray), slot_lds[k]) : make_int2(-1, -1);
            }
            if (!fits) {
                rec_ok = false;
                continue;
            }
            if (lane == 0)
                rs->ray_info[ray] = make_float4(__int_as_float(first), __int_as_float(n_rec_chunks),
                                                __expf(log_t) * o.bg, 0.f);
        } else if (lane == 0) {
            rs->ray_info[ray] = make_float4(0.f, 0.f, __expf(log_t) * o.bg, 0.f);
        }
        break;
    }
    const float tb = __expf(log_t) * o.bg;
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = total[a] + tb;
    // A replayed backward works per record and per recorded chunk, a
    // re-marching one per chunk.
    if (cost != nullptr)
        *cost = backward_cost(RECORD ? n_live_all + 4u * n_live_chunks : n_chunks);
    if (RECORD) {
        // A long ray lists its segments m >= 1 for the replay (one reservation).
        int n_seg = 1;
        if (rec_ok && n_rec_chunks >= kRecSplitMin)
            n_seg = min((n_rec_chunks + kRecSegChunks - 1) / kRecSegChunks, kRecSegs);
        const int part = static_cast<int>(ray % rs->parts);
        int local = 0;
        if (lane == 0 && n_seg > 1)
            local = static_cast<int>(atomicAdd(rs->next_block + kRecParts + part,
                                               static_cast<uint32_t>(n_seg - 1)));
        local = __shfl(local, 0, kGroup);
        const bool listed = n_seg > 1 && static_cast<int64_t>(local) + n_seg - 1 <= rs->spart;
        for (int m = 1 + lane; m < n_seg; m += kGroup) {
            const int64_t l = static_cast<int64_t>(local) + m - 1;
            if (l < rs->spart)
                rs->seg_list[part * rs->spart + l] =
                    listed ? make_int2(static_cast<int>(ray), m) : make_int2(-1, -1);
        }
        if (lane == 0) {
            if (rec_ok) put_header(rs->rec, cur, -1, 0);
            rs->state[ray] = rec_ok ? 1 : 2;
            for (int m = 1; m < kRecSegs; ++m)
                if (!rec_ok || m * kRecSegChunks >= n_rec_chunks)
                    rs->seg[ray * (kRecSegs - 1) + m - 1] = -1;
            rs->split[ray] = listed ? 1 : -1;
        }
    }
    if (LOG) {
        const int n_logged = static_cast<int>(min(n_chunks, static_cast<uint32_t>(kLogChunks)));
        wave_lds_fence();
        for (int ci = 0; ci < n_logged; ++ci) {
            const float* src = log_lds + ci * kLogWords;
            float* dst = log->rec + (ray * kLogChunks + ci) * kLogWords;
            dst[lane] = src[lane];
            if (lane < 2) dst[32 + lane] = src[32 + lane];
        }
        if (lane == 0)
            log->head[ray] = make_int4(n_logged, log_more ? 1 : 0, __float_as_int(log_next_t), 0);
    }
    if (STATS && lane == 0) {
        fnv.mix(__float_as_uint(log_t));
        stats[0] = static_cast<int32_t>(n_iter);
        stats[1] = static_cast<int32_t>(n_eval);
        stats[2] = static_cast<int32_t>(n_hit);
        stats[3] = static_cast<int32_t>(fnv.h);
    }
    return log_t;
}

struct GradOut {
    float* __restrict__ gd;   // grad density, rows gds elements apart
    float* __restrict__ gs;   // grad sh, rows gss elements apart
    uint8_t* __restrict__ mask;
    int64_t gds;
    int64_t gss;
};

// Per-sample record a backward chunk stages in LDS for its scatter loop: every
// lane of the group then reads the current sample's record at one address (a
// broadcast read) instead of gathering each field from the sample's lane.
//
// Corner quantities are kept in parity slots: grid point (X, Y, Z) lives in
// slot (X & 1) << 2 | (Y & 1) << 1 | (Z & 1).  The 8 corners of a cell take
// the 8 distinct slots, so corner k (bits dx dy dz) of the cell with lower
// corner c sits in slot k ^ parity(c).  A step to an adjacent cell keeps the
// shared grid points in their slots, and the grid points it enters take the
// slots of those it leaves: no accumulator moves, only the departing slots
// are flushed.  `w` and `lk` are stored in slot order; `depart` has bit s set
// when the grid point in slot s before this sample is not a corner of its
// cell (all 8 after a jump or for the ray's first live sample).
struct __align__(16) SampleRec {
    float w[8];    // trilinear weight of the corner in slot s
    float val[4];  // colour factors R, G, B and the density gradient
    int lk[8];     // link of the corner in slot s
    int depart;
    int pad[3];
};

// LDS of one backward workgroup: the sample records of its rays.
struct BwdShared {
    SampleRec rec[kBlock];
};

// Gradient scatter state of one ray.  Lane L of the 32-lane group owns one
// gradient column of a row: SH channel L (c = L / B, j = L % B) for L < C,
// the density for L == C; lanes above are idle.  For each parity slot s the
// lane keeps in registers the accumulated factor of its quantity at the
// slot's grid point - the colour factor of channel c (the SH gradient is that
// times basis_j) or the density gradient - and the grid point's link.
// Consecutive samples add into the same accumulators while their cells share
// the grid point, so a grid point is flushed once per visit.  A flush of slot
// s is one no-return atomic instruction over lanes 0..C: the row's SH
// channels and its density, the same 128-byte row in the interleaved layout.
// (The touched-row mask is marked when a sample enters a grid point, while
// the records are built.)
template <int B>
struct Scatter {
    static constexpr int C = 3 * B;
    float acc[8];
    int link[8];
    int c0 = 0, c1 = 0, c2 = 0;  // current cell (when `have`)
    bool have = false;           // a cell has been entered
    char* col = nullptr;         // this lane's gradient column (null: idle lane)
    uint32_t col_stride = 0;     // bytes between the column's rows
    float mult = 0.f;            // basis_j for SH lanes, 1 for the density lane

    __device__ __forceinline__ void init(const GradOut& go, int lane, float basis_j) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            acc[k] = 0.f;
            link[k] = -1;
        }
        if (lane < C && go.gs != nullptr) {
            col = reinterpret_cast<char*>(go.gs + lane);
            col_stride = static_cast<uint32_t>(go.gss * 4);
            mult = basis_j;
        } else if (lane == C && go.gd != nullptr) {
            col = reinterpret_cast<char*>(go.gd);
            col_stride = static_cast<uint32_t>(go.gds * 4);
            mult = 1.f;
        }
    }

    // Adds the accumulated factor a of the grid point with link lk to its row.
    template <typename DT>
    __device__ __forceinline__ void emit(const GridView<DT>& g, int lk, float a) const {
        if (col != nullptr && g.valid(lk))
            atomic_add_f32(reinterpret_cast<float*>(
                               col + static_cast<uint64_t>(static_cast<uint32_t>(lk)) * col_stride),
                           mult * a);
    }

    // Flushes the slots of `out` (bit s: slot s) and clears them.
    template <typename DT>
    __device__ __forceinline__ void flush(const GridView<DT>& g, int out) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool go = (out >> k) & 1;
            emit(g, go ? link[k] : -1, acc[k]);
            acc[k] = go ? 0.f : acc[k];
        }
    }

    // Flushes all 8 slots and clears them.
    template <typename DT>
    __device__ __forceinline__ void flush_all(const GridView<DT>& g, const GradOut&, int) {
        flush(g, 0xFF);
#pragma unroll
        for (int k = 0; k < 8; ++k) link[k] = -1;
    }
};

// Backward state of one ray: dL/drgb (gq), the running accum = rgb . gq (+
// beta term) of the reference, and the corner accumulators.
template <int B>
struct BwdRay {
    static constexpr int C = 3 * B;
    float gq[3];
    float accum;
    float sparsity;
    int my_q;        // quantity of this lane's column: c for SH lanes, 3 otherwise
    SampleRec* rec;  // this group's LDS records
    Scatter<B> sc;

    __device__ __forceinline__ BwdRay(const RayCtx<B>& rc, int lane, const float* g3, float acc0,
                                      float sp, BwdShared* sh, const GradOut& go) {
        const int gbase = threadIdx.x & ~(kGroup - 1);
#pragma unroll
        for (int a = 0; a < 3; ++a) gq[a] = g3[a];
        accum = acc0;
        sparsity = sp;
        // Selects, not a register array indexed by lane (that goes to scratch).
        float basis_j = 0.f;
#pragma unroll
        for (int j = 0; j < B; ++j) basis_j = (lane < C && lane % B == j) ? rc.basis[j] : basis_j;
        my_q = lane < C ? lane / B : 3;
        rec = sh->rec + gbase;
        sc.init(go, lane, basis_j);
    }
};

// One chunk's contribution to the gradients (trace_ray_cuvol_backward,
// render_lerp_kernel_cuvol.cu:227-383).  Lane i holds sample i of the chunk if
// live: its position (p, l), corner links, colour + 0.5 (cs), weight, log
// transmittance after it, world step and density.  The gradient scalars are
// formed sample-parallel - the colour factors weight * [colour >= 0] * g and
// the density gradient ws * (total * T_after - accum), with the running
// `accum` as a prefix sum over the chunk - and staged as LDS records with each
// sample's 8 trilinear weights and links in parity-slot order and the slots
// its step leaves; then the samples are added in order into the slot
// accumulators, one fma per corner each.
template <typename DT, int B>
__device__ __forceinline__ void scatter_chunk(const GridView<DT>& g, const GradOut& go,
                                              BwdRay<B>& st, int lane, bool live,
                                              uint32_t live_mask, const float* p, const int* l,
                                              const int* lk, const float* cs, float weight,
                                              float lt_after, float ws, float sigma) {
    float total = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) total += fmaxf(cs[a], 0.f) * st.gq[a];
    const float run = group_scan(live ? weight * total : 0.f);
    const float acc_s = st.accum - run;
    st.accum -= group_lane(run, 31);
    float gsig = ws * (total * __expf(lt_after) - acc_s);
    if (st.sparsity > 0.f) gsig += st.sparsity * (4.f * sigma / (1.f + 2.f * (sigma * sigma)));
    Scatter<B>& sc = st.sc;
    // Cell of the previous live sample: an earlier lane of this chunk, or the
    // cell the scatter is in.
    const uint32_t before = live_mask & ((1u << lane) - 1u);
    const int prev = before ? 31 - __clz(before) : 0;
    int pl[3];
    pl[0] = __shfl(l[0], prev, kGroup);
    pl[1] = __shfl(l[1], prev, kGroup);
    pl[2] = __shfl(l[2], prev, kGroup);
    if (!before) {
        pl[0] = sc.c0;
        pl[1] = sc.c1;
        pl[2] = sc.c2;
    }
    const bool had = before != 0u || sc.have;
    if (live) {
        const int d0 = l[0] - pl[0], d1 = l[1] - pl[1], d2 = l[2] - pl[2];
        // Slots the step from the previous live sample's cell leaves: along an
        // axis stepped by +1 the grid points at the old lower coordinate, by
        // -1 those at the old upper one; their slots have that coordinate's
        // parity in the axis' bit.
        int depart;
        if (!had || d0 < -1 || d0 > 1 || d1 < -1 || d1 > 1 || d2 < -1 || d2 > 1) {
            depart = 0xFF;
        } else {
            const int bx = (pl[0] & 1) ^ (d0 < 0 ? 1 : 0);
            const int by = (pl[1] & 1) ^ (d1 < 0 ? 1 : 0);
            const int bz = (pl[2] & 1) ^ (d2 < 0 ? 1 : 0);
            depart = (d0 != 0 ? (bx ? 0xF0 : 0x0F) : 0) | (d1 != 0 ? (by ? 0xCC : 0x33) : 0) |
                     (d2 != 0 ? (bz ? 0xAA : 0x55) : 0);
        }
        SampleRec& R = st.rec[lane];
        // Slot s holds corner s ^ parity(l): swapping each axis' pair of
        // weights (and links) by the cell's parity puts them in slot order.
        const bool px = l[0] & 1, py = l[1] & 1, pz = l[2] & 1;
        const float wx[2] = {px ? p[0] : 1.f - p[0], px ? 1.f - p[0] : p[0]};
        const float wy[2] = {py ? p[1] : 1.f - p[1], py ? 1.f - p[1] : p[1]};
        const float wz[2] = {pz ? p[2] : 1.f - p[2], pz ? 1.f - p[2] : p[2]};
#pragma unroll
        for (int k = 0; k < 8; ++k) R.w[k] = wy[(k >> 1) & 1] * wz[k & 1] * wx[k >> 2];
        int q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = lk[k];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k & 4) continue;
            const int a0 = q[k], a1 = q[k | 4];
            q[k] = px ? a1 : a0;
            q[k | 4] = px ? a0 : a1;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k & 2) continue;
            const int a0 = q[k], a1 = q[k | 2];
            q[k] = py ? a1 : a0;
            q[k | 2] = py ? a0 : a1;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k & 1) continue;
            const int a0 = q[k], a1 = q[k | 1];
            q[k] = pz ? a1 : a0;
            q[k | 1] = pz ? a0 : a1;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
            R.val[a] = weight * ((fmaxf(cs[a], 0.f) == cs[a]) ? 1.f : 0.f) * st.gq[a];
        R.depart = depart;
        R.val[3] = gsig;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) R.lk[kk] = q[kk];
        // Each visit of a grid point is entered by exactly one sample: mark
        // its row there (the scatter flushes it once when the visit ends).
        if (go.mask != nullptr) {
#pragma unroll
            for (int kk = 0; kk < 8; ++kk)
                if (((depart >> kk) & 1) && g.valid(q[kk])) go.mask[q[kk]] = 1;
        }
    }
    // The scatter ends the chunk in the last live sample's cell.
    const int last = 31 - __clz(live_mask);
    sc.c0 = group_bcast(l[0], last);
    sc.c1 = group_bcast(l[1], last);
    sc.c2 = group_bcast(l[2], last);
    sc.have = true;
    wave_lds_fence();
    // Two staged records alternate: while one sample is added, the next
    // one's record is on its way from LDS.  Every read is unconditional
    // (a lane past the last live sample reads the first one again), so each
    // wait covers exactly the record it needs.
    struct Staged {
        float4 wa, wb;
        int4 la, lb;
        float v;
        int depart;
    };
    auto stage = [&](int s) {
        const SampleRec& R = st.rec[s];
        Staged q;
        q.wa = *reinterpret_cast<const float4*>(R.w);
        q.wb = *reinterpret_cast<const float4*>(R.w + 4);
        q.la = *reinterpret_cast<const int4*>(R.lk);
        q.lb = *reinterpret_cast<const int4*>(R.lk + 4);
        q.v = R.val[st.my_q];
        q.depart = R.depart;
        return q;
    };
    auto add = [&](const Staged& q) {
        if (q.depart != 0) {
            sc.flush(g, q.depart);
            sc.link[0] = q.la.x; sc.link[1] = q.la.y; sc.link[2] = q.la.z; sc.link[3] = q.la.w;
            sc.link[4] = q.lb.x; sc.link[5] = q.lb.y; sc.link[6] = q.lb.z; sc.link[7] = q.lb.w;
        }
        const float w[8] = {q.wa.x, q.wa.y, q.wa.z, q.wa.w, q.wb.x, q.wb.y, q.wb.z, q.wb.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) sc.acc[k] = fmaf(w[k], q.v, sc.acc[k]);
    };
    uint32_t m = live_mask;
    const int s0 = __ffs(m) - 1;
    m &= m - 1u;
    Staged x = stage(s0);
    bool have_y = m != 0u;
    Staged y = stage(have_y ? __ffs(m) - 1 : s0);
    m &= m - 1u;
    while (true) {
        add(x);
        if (!have_y) break;
        const bool have_x = m != 0u;
        x = stage(have_x ? __ffs(m) - 1 : s0);
        m &= m - 1u;
        add(y);
        if (!have_x) break;
        have_y = m != 0u;
        y = stage(have_y ? __ffs(m) - 1 : s0);
        m &= m - 1u;
    }
    wave_lds_fence();  // records are rewritten by the next chunk
}

// Backward of one ray by re-marching it: every chunk is rematerialised
// exactly as in the forward, its colours gathered again.
template <typename DT, int B, bool PREFETCH = true>
__device__ __forceinline__ void march_backward(const GridView<DT>& g, const OptView& o,
                                               const Ray& r, const RayCtx<B>& rc, int lane,
                                               BwdRay<B>& st, const GradOut& go,
                                               ChunkLog log_v = ChunkLog{},
                                               int64_t ray = 0) {
    if (r.tmin > r.tmax) return;
    // Chunks the forward logged are evaluated from their placement.
    int n_logged = 0;
    bool log_more = false;
    float log_next_t = 0.f;
    if (log_v.rec != nullptr) {
        const int4 h = log_v.head[ray];
        n_logged = min(h.x, kLogChunks);
        log_more = h.y != 0;
        log_next_t = __int_as_float(h.z);
    }
    // With PREFETCH the next logged chunk's placement is read, and its corner
    // links requested, while the current chunk is evaluated, so a logged
    // chunk starts with its links on the way (it needs the registers of a
    // 3-wave build).
    float nt = 0.f;
    int nplace = 0;
    uint32_t nmask = 0u;
    int nlk[8];
    if (PREFETCH && n_logged > 0) {
        const float* rec = log_v.rec + ray * kLogChunks * kLogWords;
        nt = rec[lane];
        nplace = __float_as_int(rec[32]);
        nmask = __float_as_uint(rec[33]);
        float np[3];
        int nl[3];
        load_corner_links(g, lane, sample_cell(g, r, nt, np, nl), nmask, nlk);
    }
    float t = r.tmin;
    float log_t = 0.f;
    int guard = 0;
    bool more = true;
    for (int ci = 0; more && ++guard < kMaxIters; ++ci) {
        Chunk c;
        if (ci < n_logged && !PREFETCH) {
            const float* rec = log_v.rec + (ray * kLogChunks + ci) * kLogWords;
            c.t = rec[lane];
            c.n_place = __float_as_int(rec[32]);
            c.load_mask = __float_as_uint(rec[33]);
            c.base = sample_cell(g, r, c.t, c.p, c.l);
            Placement pl;
            pl.more = ci + 1 < n_logged || log_more;
            pl.next_t = log_next_t;  // read only after the last logged chunk
            eval_chunk(g, o, r, lane, log_t, pl, c);
        } else if (ci < n_logged) {
            c.t = nt;
            c.n_place = nplace;
            c.load_mask = nmask;
            c.base = sample_cell(g, r, c.t, c.p, c.l);
#pragma unroll
            for (int k = 0; k < 8; ++k) c.lk[k] = nlk[k];
            const bool next = ci + 1 < n_logged;
            if (next) {
                const float* rec = log_v.rec + (ray * kLogChunks + ci + 1) * kLogWords;
                nt = rec[lane];
                nplace = __float_as_int(rec[32]);
                nmask = __float_as_uint(rec[33]);
            }
            Placement pl;
            pl.more = next || log_more;
            pl.next_t = log_next_t;  // read only after the last logged chunk
            eval_chunk(g, o, r, lane, log_t, pl, c, true);
            if (next && c.more) {
                float np[3];
                int nl[3];
                load_corner_links(g, lane, sample_cell(g, r, nt, np, nl), nmask, nlk);
            }
        } else {
            march_chunk(g, o, r, lane, t, log_t, c);
        }
        if (c.live_mask != 0u) {
            float cs[3] = {0.f, 0.f, 0.f};
            chunk_colours<DT, B, true>(g, rc, c, lane, cs, nullptr);
            scatter_chunk<DT, B>(g, go, st, lane, c.live, c.live_mask, c.p, c.l, c.lk, cs,
                                 c.weight, c.lt_after, c.ws, c.sigma);
        }
        log_t = c.log_t_end;
        t = c.next_t;
        more = c.more;
    }
    st.sc.flush_all(g, go, lane);
}

// Backward of one ray from its record stream (the fused call's forward wrote
// every contributing sample with its colour and transmittance): no link,
// density or SH reads, only the records and the gradient scatter.  Starts at
// the header slot pos and replays at most max_chunks recorded chunks.
template <typename DT, int B>
__device__ __forceinline__ void replay_backward(const GridView<DT>& g, const RecStream& rs,
                                                int64_t pos, int max_chunks, int lane,
                                                BwdRay<B>& st, const GradOut& go) {
    int chunks = 0;
    for (int guard = 0; guard < kMaxIters; ++guard) {
        const float4 h = rs.rec[6 * pos];
        const int n = __float_as_int(h.x);
        if (n < 0) break;
        if (n == 0) {
            pos = __float_as_int(h.y);
            continue;
        }
        if (chunks++ == max_chunks) break;  // the next segment's first chunk
        const bool live = lane < n;
        float p[3] = {0.f, 0.f, 0.f}, cs[3] = {0.f, 0.f, 0.f};
        int l[3] = {0, 0, 0}, lk[8];
        float weight = 0.f, lt_after = 0.f, ws = 0.f, sigma = 0.f;
        if (live) {
            const float4* q = rec_field0(rs.rec, pos) + lane;
            const float4 a = q[0], b = q[n], c2 = q[2 * n], d = q[3 * n], e = q[4 * n], f = q[5 * n];
            p[0] = a.x; p[1] = a.y; p[2] = a.z;
            l[0] = __float_as_int(b.x); l[1] = __float_as_int(b.y); l[2] = __float_as_int(b.z);
            weight = b.w;
            lk[0] = __float_as_int(c2.x); lk[1] = __float_as_int(c2.y);
            lk[2] = __float_as_int(c2.z); lk[3] = __float_as_int(c2.w);
            lk[4] = __float_as_int(d.x); lk[5] = __float_as_int(d.y);
            lk[6] = __float_as_int(d.z); lk[7] = __float_as_int(d.w);
            cs[0] = e.x; cs[1] = e.y; cs[2] = e.z; lt_after = e.w;
            ws = f.x; sigma = f.y;
        } else {
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) lk[kk] = -1;
        }
        const uint32_t live_mask = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
        scatter_chunk<DT, B>(g, go, st, lane, live, live_mask, p, l, lk, cs, weight, lt_after,
                             ws, sigma);
        pos += n + 1;
    }
    st.sc.flush_all(g, go, lane);
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Minimum waves per SIMD the march kernels are compiled for (the register
// budget is 512 / waves VGPRs).  The forward is latency bound: occupancy
// hides its dependent link -> row loads.  The fused backward is compiled
// twice: at 3 waves, whose 170 registers hold the next chunk's prefetched
// placement and links without spilling (fastest while a batch fills the
// device a few times over), and at 4 waves for larger batches, where
// throughput counts more than each ray's latency.
#ifndef SVOX2_FWD_WAVES
#define SVOX2_FWD_WAVES 5
#endif
#ifndef SVOX2_BWD_WAVES
#define SVOX2_BWD_WAVES 4
#endif
#ifndef SVOX2_BWD_WAVES_FEW
#define SVOX2_BWD_WAVES_FEW 3
#endif
#define SVOX2_FWD_KERNEL __global__ __launch_bounds__(kBlock) \
    __attribute__((amdgpu_waves_per_eu(SVOX2_FWD_WAVES)))
#define SVOX2_BWD_KERNEL __global__ __launch_bounds__(kBlock) \
    __attribute__((amdgpu_waves_per_eu(SVOX2_BWD_WAVES)))
// Record the fused call's samples at every batch size (a tuning switch; by
// default only batches below the ordering threshold record).
#ifndef SVOX2_RECORD_ALL
#define SVOX2_RECORD_ALL 0
#endif
#ifndef SVOX2_REC_WAVES
#define SVOX2_REC_WAVES 4
#endif
// The replay kernel at 5 waves: 96 registers hold the scatter loop's two
// staged records without spilling.
#ifndef SVOX2_REPLAY_WAVES
#define SVOX2_REPLAY_WAVES 5
#endif
#define SVOX2_REPLAY_KERNEL __global__ __launch_bounds__(kBlock) \
    __attribute__((amdgpu_waves_per_eu(SVOX2_REPLAY_WAVES)))
#define SVOX2_REC_KERNEL __global__ __launch_bounds__(kBlock) \
    __attribute__((amdgpu_waves_per_eu(SVOX2_REC_WAVES)))

// The recording forward of the fused call: a kernel of its own so that it is
// compiled for the backward's register budget (the record writes and the
// colour-to-lane moves need more registers than the plain forward has).
template <typename DT, int B>
SVOX2_REC_KERNEL void k_render_rays_rec(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        float* __restrict__ rgb_out, float* __restrict__ log_t_out,
        uint32_t* __restrict__ cost_out, RecStream rs, const uint32_t* __restrict__ order) {
    const int64_t slot = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (slot >= n) return;
    const int64_t ray = order != nullptr ? static_cast<int64_t>(order[slot]) : slot;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    float out[3];
    uint32_t cost = 0;
    // LDS staging of this group's recorded chunk header slots.
    __shared__ int32_t slot_lds[(kBlock / kGroup) * kDeferChunks];
    const float log_t = march_forward<DT, B, false, true, false, true>(
        g, o, r, rc, lane, out, nullptr, cost_out != nullptr ? &cost : nullptr, &rs, ray,
        nullptr, nullptr, slot_lds + (threadIdx.x / kGroup) * kDeferChunks);
    // A ray whose chunks were listed gets its colour from k_rec_colours.
    if (lane < 3) rgb_out[3 * ray + lane] = pick3(out, lane);
    if (log_t_out != nullptr && lane == 0) log_t_out[ray] = log_t;
    if (cost_out != nullptr && lane == 0) cost_out[ray] = cost;
}

// Colours of the chunks the recording forward listed, one group per chunk
// (the chunks of a ray run on different groups at once).  A group reads the
// chunk's records (positions, cells, corner links, weights: its live samples
// compacted into lanes 0 .. n - 1, which gives the forward's runs), forms the
// colours as the forward does, writes each sample's colour + 0.5 into its
// record and the chunk's colour into chunk_rgb.  The replay sums a ray's
// chunk colours in march order (no group here waits for another).
template <typename DT, int B>
__global__ __launch_bounds__(kBlock) void k_rec_colours(GridView<DT> g,
                                                        const float* __restrict__ dirs,
                                                        RecStream rs) {
    const int lane = threadIdx.x % kGroup;
    const int64_t gid = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int64_t ngroups = static_cast<int64_t>(gridDim.x) * (kBlock / kGroup);
    const int64_t total = rs.parts * rs.cpart;
    for (int64_t j = gid; j < total; j += ngroups) {
        const int64_t part = j / rs.cpart;
        if (j - part * rs.cpart >= min(static_cast<int64_t>(rs.next_block[2 * kRecParts + part]),
                                       rs.cpart))
            continue;  // past the partition's entries
        const int2 e = rs.clist[j];
        if (e.x < 0) continue;  // entries of a ray whose list did not fit
        const int64_t ray = e.x, slot = e.y;
        const int n = __float_as_int(rs.rec[6 * slot].x);
        Chunk c;
        c.live = lane < n;
        c.live_mask = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
        float4* q = rec_field0(rs.rec, slot) + lane;
        if (c.live) {
            const float4 a = q[0], b = q[n], c2 = q[2 * n], d = q[3 * n];
            c.p[0] = a.x;
            c.p[1] = a.y;
            c.p[2] = a.z;
            c.base = __float_as_int(a.w);
            c.weight = b.w;
            c.lk[0] = __float_as_int(c2.x);
            c.lk[1] = __float_as_int(c2.y);
            c.lk[2] = __float_as_int(c2.z);
            c.lk[3] = __float_as_int(c2.w);
            c.lk[4] = __float_as_int(d.x);
            c.lk[5] = __float_as_int(d.y);
            c.lk[6] = __float_as_int(d.z);
            c.lk[7] = __float_as_int(d.w);
        } else {
            c.p[0] = c.p[1] = c.p[2] = 0.f;
            c.base = -1;
            c.weight = 0.f;
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) c.lk[kk] = -1;
        }
        const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
        const RayCtx<B> rc(wd);
        float cs[3] = {0.f, 0.f, 0.f}, cacc[3] = {0.f, 0.f, 0.f}, sum[3], unused[3] = {0.f, 0.f, 0.f};
        chunk_colours<DT, B, true>(g, rc, c, lane, cs, cacc);
        add_chunk_colour(unused, cacc, sum);
        if (c.live) {
            float* q4 = reinterpret_cast<float*>(q + 4 * n);
            q4[0] = cs[0];
            q4[1] = cs[1];
            q4[2] = cs[2];
        }
        if (lane == 0) rs.chunk_rgb[j] = make_float4(sum[0], sum[1], sum[2], 0.f);
    }
}

template <typename DT, int B, bool STATS, bool LOG = false>
SVOX2_FWD_KERNEL void k_render_rays(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        float* __restrict__ rgb_out, float* __restrict__ log_t_out,
        int32_t* __restrict__ stats, uint32_t* __restrict__ cost_out, RecStream rs,
        const uint32_t* __restrict__ order, ChunkLog log) {
    const int64_t slot = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (slot >= n) return;
    const int64_t ray = order != nullptr ? static_cast<int64_t>(order[slot]) : slot;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    float out[3];
    uint32_t cost = 0;
    // LDS staging of this group's chunk log (used when LOG).
    __shared__ float log_lds[LOG ? (kBlock / kGroup) * kLogChunks * kLogWords : 1];
    const float log_t = march_forward<DT, B, STATS, false, LOG>(
        g, o, r, rc, lane, out, STATS ? stats + 4 * ray : nullptr,
        cost_out != nullptr ? &cost : nullptr, &rs, ray, &log,
        log_lds + (threadIdx.x / kGroup) * kLogChunks * kLogWords);
    if (!STATS) {
        if (lane < 3) rgb_out[3 * ray + lane] = pick3(out, lane);
        if (log_t_out != nullptr && lane == 0) log_t_out[ray] = log_t;
        if (cost_out != nullptr && lane == 0) cost_out[ray] = cost;
    }
}

// Depth along each ray (trace_ray_expected_term / trace_ray_sigma_thresh,
// render_lerp_kernel_cuvol.cu:122-225).  Both walk the colour march's
// samples: the same placement, skips, densities and (for the expected term)
// transmittance and stop, so only the accumulator differs.
//   kDepthExpected: sum over contributing samples of w (t / step) ws, the
//     expected termination distance along the unit world direction;
//   kDepthSigma: (t / step) ws of the first sample whose density exceeds the
//     given threshold (no transmittance or stop; the host passes that
//     threshold as o.sigma_thresh and a log_stop of -inf).
// A ray that misses the grid gets 0.  Within a chunk the terms are summed
// across lanes, then chunk sums in march order.
constexpr int kDepthExpected = 0;
constexpr int kDepthSigma = 1;

template <typename DT, int MODE>
SVOX2_FWD_KERNEL void k_render_depth(GridView<DT> g, OptView o,
                                     const float* __restrict__ origins,
                                     const float* __restrict__ dirs,
                                     const float* __restrict__ clip, int64_t n,
                                     float* __restrict__ depth_out) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (ray >= n) return;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    float depth = 0.f;
    if (r.tmin <= r.tmax) {
        float t = r.tmin, log_t = 0.f;
        bool more = true;
        int guard = 0;
        while (more && ++guard < kMaxIters) {
            Chunk c;
            march_chunk(g, o, r, lane, t, log_t, c);
            if (MODE == kDepthSigma) {
                if (c.live_mask != 0) {
                    const float tj = group_bcast(c.t, __ffs(c.live_mask) - 1);
                    depth = (tj / o.step) * r.ws;
                    break;
                }
            } else if (c.live_mask != 0) {
                float term = c.live ? c.weight * (c.t / o.step) * r.ws : 0.f;
                term += __shfl_xor(term, 1, kGroup);
                term += __shfl_xor(term, 2, kGroup);
                term += __shfl_xor(term, 4, kGroup);
                term += __shfl_xor(term, 8, kGroup);
                term += __shfl_xor(term, 16, kGroup);
                depth += term;
            }
            log_t = c.log_t_end;
            t = c.next_t;
            more = c.more;
        }
    }
    if (lane == 0) depth_out[ray] = depth;
}

// ---------------------------------------------------------------------------
// The nvol backend (render_lerp_kernel_nvol.cu): the colour march's sample
// placement and densities, composited by accumulated opacity instead of
// transmittance.  Each contributing sample (sigma > sigma_thresh) adds
// 1 - exp(-ws sigma) to the ray's opacity, A = min(A + 1 - exp(-ws sigma), 1)
// (:84-88), and weighs its colour by the increase of A; the ray stops at the
// sample where A reaches 1 (:93); the background gets (1 - A) bg (:98-100).
// There is no transmittance stop and no opaque last sample.
//
// The backward follows the reference's two passes (trace_ray_nvol_backward,
// :109-292).  Pass 1, per contributing sample up to the saturating one: the
// colour terms weight * [colour >= 0] * g_c * basis_j into the SH rows, and,
// unless the sample saturated A, ws exp(-ws sigma) * sum_c g_c max(colour_c,
// 0) (plus the sparsity term) into the densities.  The saturating sample's
// sum_c g_c max(colour_c, 0), or bg * (g_0 + g_1 + g_2) when A never reached
// 1, is the "last colour" L.  Pass 2, if L != 0: every evaluated sample up to
// the saturating one, whatever its density, accumulates its own opacity the
// same way, stops at the one that reaches 1, and adds -ws exp(-ws sigma) L to
// the densities.  Density terms mark their corners' rows in the mask.
// ---------------------------------------------------------------------------

// Opacity accumulation over a chunk's samples in lane order, one per lane
// where acc is set (e = exp(-ws sigma) of its sample): returns the lane's
// weight (the increase of alpha at its sample) and sets sat to the first lane
// whose sample brings alpha to 1 (32 if none).  alpha is the group's running
// opacity, the same in every lane of the group.
__device__ __forceinline__ float nvol_accumulate(int lane, bool acc, float e, float& alpha,
                                                 int& sat) {
    // Every accumulated increase 1 - e non-negative (e <= 1) in the whole
    // wavefront: min(A + x, 1) applied in lane order is min(A + prefix sum,
    // 1), so one inclusive scan replaces the serial walk.
    if (__ballot(acc && e > 1.f) == 0ull) {
        const float a = acc ? 1.f - e : 0.f;
        const float incl = group_scan(a);
        const float after = fminf(alpha + incl, 1.f);
        const float before = fminf(alpha + (incl - a), 1.f);
        const uint32_t sm = group_ballot(acc && after >= 1.f);
        sat = sm ? __ffs(sm) - 1 : 32;
        alpha = fminf(alpha + group_lane(incl, 31), 1.f);
        return acc && lane <= sat ? after - before : 0.f;
    }
    const uint32_t m = group_ballot(acc);
    const int n_rep = wave_last_set(acc);
    float my_w = 0.f;
    sat = 32;
    for (int j = 0; j < n_rep; ++j) {
        const float ej = group_lane(e, j);
        if (((m >> j) & 1u) && sat == 32) {
            const float na = fminf(alpha + 1.f - ej, 1.f);
            if (lane == j) my_w = na - alpha;
            alpha = na;
            if (alpha >= 1.f) sat = j;
        }
    }
    return my_w;
}

// nvol forward of one ray per 32-lane group: rgb_out [n][3].
template <typename DT, int B>
SVOX2_FWD_KERNEL void k_render_nvol(GridView<DT> g, OptView o, const float* __restrict__ origins,
                                    const float* __restrict__ dirs,
                                    const float* __restrict__ clip, int64_t n,
                                    float* __restrict__ rgb_out) {
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (ray >= n) return;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    float alpha = 0.f, total[3] = {0.f, 0.f, 0.f};
    if (r.tmin <= r.tmax) {
        float t = r.tmin;
        bool more = true;
        int guard = 0;
        while (more && ++guard < kMaxIters) {
            Chunk c;
            march_chunk(g, o, r, lane, t, 0.f, c);
            const float e = c.contrib ? __expf(-c.ws * c.sigma) : 1.f;
            int sat;
            c.weight = nvol_accumulate(lane, c.contrib, e, alpha, sat);
            c.live = c.contrib && lane <= sat;
            c.live_mask = group_ballot(c.live);
            if (c.live_mask != 0u) {
                float cs[3], cacc[3] = {0.f, 0.f, 0.f}, sum[3];
                chunk_colours<DT, B, false>(g, rc, c, lane, cs, cacc);
                add_chunk_colour(total, cacc, sum);
            }
            if (sat < 32) break;
            t = c.next_t;
            more = c.more;
        }
    }
    if (lane < 3) rgb_out[3 * ray + lane] = pick3(total, lane) + (1.f - alpha) * o.bg;
}

// Adds v * (trilinear weight of corner k at p) into column col of the rows
// of the corners with a row.
__device__ __forceinline__ void nvol_scatter(float* base, int64_t stride, int col, const int* lk,
                                             const float* p, float v, int rows) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (static_cast<uint32_t>(lk[k]) < static_cast<uint32_t>(rows))
            atomic_add_f32(base + lk[k] * stride + col, corner_weight(k, p) * v);
}

// nvol backward of one ray per 32-lane group.  grad_in: dL/drgb [n][3], or
// with rgb_residual the targets (the fused MSE step: g = (color_cache -
// grad_in) * norm).
template <typename DT, int B>
__global__ __launch_bounds__(kBlock) void k_render_nvol_backward(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        const float* __restrict__ grad_in, const float* __restrict__ color_cache,
        int rgb_residual, float norm, float sparsity, GradOut go) {
    constexpr int C = 3 * B;
    const int64_t ray = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (ray >= n) return;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    if (r.tmin > r.tmax) return;
    float gq[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        gq[a] = rgb_residual ? (color_cache[3 * ray + a] - grad_in[3 * ray + a]) * norm
                             : grad_in[3 * ray + a];
    // This lane's SH coefficient (lane < C): its channel and basis value.
    float basis_j = 0.f;
#pragma unroll
    for (int j = 0; j < B; ++j) basis_j = (lane < C && lane % B == j) ? rc.basis[j] : basis_j;
    const int my_c = lane < C ? lane / B : 0;

    // Pass 1.
    float alpha = 0.f, last = 0.f, t_sat = r.tmax;
    bool saturated = false;
    {
        float t = r.tmin;
        bool more = true;
        int guard = 0;
        while (more && ++guard < kMaxIters) {
            Chunk c;
            march_chunk(g, o, r, lane, t, 0.f, c);
            const float e = c.contrib ? __expf(-c.ws * c.sigma) : 1.f;
            int sat;
            c.weight = nvol_accumulate(lane, c.contrib, e, alpha, sat);
            c.live = c.contrib && lane <= sat;
            c.live_mask = group_ballot(c.live);
            float cs[3] = {0.f, 0.f, 0.f};
            if (c.live_mask != 0u) chunk_colours<DT, B, true>(g, rc, c, lane, cs, nullptr);
            // The sample's gradient scalars (live lanes).
            float gc[3] = {0.f, 0.f, 0.f}, tot = 0.f;
            if (c.live) {
                float tc[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float cl = fmaxf(cs[a], 0.f);
                    tc[a] = cl * gq[a];
                    gc[a] = c.weight * (cl == cs[a] ? 1.f : 0.f) * gq[a];
                }
                tot = (tc[0] + tc[2]) + tc[1];
                if (lane != sat) {  // not the saturating sample: its density term
                    float gs = c.ws * e * tot;
                    if (sparsity > 0.f)
                        gs += sparsity * (4.f * c.sigma / (1.f + 2.f * (c.sigma * c.sigma)));
                    if (go.gd) nvol_scatter(go.gd, go.gds, 0, c.lk, c.p, gs, g.rows);
                    if (go.mask) {
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                            if (static_cast<uint32_t>(c.lk[k]) < static_cast<uint32_t>(g.rows))
                                go.mask[c.lk[k]] = 1;
                    }
                }
            }
            // SH terms, sample by sample: lane = coefficient.
            uint32_t todo = c.live_mask;
            while (todo != 0u) {
                const int s = __ffs(todo) - 1;
                todo &= todo - 1u;
                float p[3];
                int lk[8];
#pragma unroll
                for (int a = 0; a < 3; ++a) p[a] = group_bcast(c.p[a], s);
#pragma unroll
                for (int k = 0; k < 8; ++k) lk[k] = group_bcast(c.lk[k], s);
                float v = group_bcast(gc[0], s);
                const float g1 = group_bcast(gc[1], s), g2 = group_bcast(gc[2], s);
                v = my_c == 0 ? v : (my_c == 1 ? g1 : g2);
                if (lane < C && v != 0.f && go.gs)
                    nvol_scatter(go.gs, go.gss, lane, lk, p, basis_j * v, g.rows);
            }
            if (sat < 32) {
                last = group_bcast(tot, sat);
                t_sat = group_bcast(c.t, sat);
                saturated = true;
                break;
            }
            t = c.next_t;
            more = c.more;
        }
    }
    if (!saturated) last = o.bg * ((gq[0] + gq[1]) + gq[2]);
    if (last == 0.f || go.gd == nullptr) return;

    // Pass 2: every evaluated sample up to the saturating one.
    Ray r2 = r;
    r2.tmax = t_sat;
    float alpha2 = 0.f;
    float t = r2.tmin;
    bool more = true;
    int guard = 0;
    while (more && ++guard < kMaxIters) {
        Chunk c;
        march_chunk(g, o, r2, lane, t, 0.f, c);
        const bool ev = lane < c.n_eval;
        const float e = ev ? __expf(-c.ws * c.sigma) : 1.f;
        int sat;
        nvol_accumulate(lane, ev, e, alpha2, sat);
        if (ev && lane < sat) {
            nvol_scatter(go.gd, go.gds, 0, c.lk, c.p, -c.ws * e * last, g.rows);
            if (go.mask) {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (static_cast<uint32_t>(c.lk[k]) < static_cast<uint32_t>(g.rows))
                        go.mask[c.lk[k]] = 1;
            }
        }
        if (sat < 32) break;
        t = c.next_t;
        more = c.more;
    }
}

// ---------------------------------------------------------------------------
// The svox1 backend (render_svox1_kernel.cu): nearest-neighbour rendering by a
// walk over whole cells.  The grid origin is shifted by +0.5 and the box is
// [0, size - 1]; from t = max(0, box entry) an occupied cell is one sample of
// its own density and colour, whose step is the distance to the cell's exit
// plus 1e-2; an empty block (link < 0: 2^(-link - 1) cells on a side, as
// accelerate() writes) is skipped to its exit plus 1e-2.  Each step depends
// on the one before, so a thread walks one ray; the float operations are in
// the reference's order (no contraction) so that cell decisions match the C
// restatement bit for bit.  Links >= rows count as empty single cells.
// ---------------------------------------------------------------------------
constexpr int kNnBlock = 256;

struct NnRay {
    float o[3], d[3], inv[3];
    float ds;       // world length per grid unit
    float t, tmax;  // walk interval (grid units)
};

// Ray set-up (trace_ray, :55-92).  skip_flat: the backward's box test, which
// leaves out axes the direction does not move along (:185-198); the forward
// applies all three.  Returns whether the ray meets the box.
template <typename DT>
__device__ __forceinline__ bool nn_setup(const GridView<DT>& g, const float* wo, const float* wd,
                                         bool skip_flat, NnRay& r) {
    float d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.o[a] = fmaf(wo[a], g.scaling[a], g.offset[a]);
        d[a] = wd[a] * g.scaling[a];
    }
    r.ds = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.d[a] = d[a] * r.ds;
        r.inv[a] = r.d[a] == 0.f ? 1e9f : static_cast<float>(1.0 / static_cast<double>(r.d[a]));
    }
    r.t = 0.f;
    r.tmax = 1e9f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.o[a] += 0.5f;
        const float t1 = (0.0f - r.o[a]) * r.inv[a];
        const float t2 = (static_cast<float>(g.size[a]) - 1.f - r.o[a]) * r.inv[a];
        if (!skip_flat || r.d[a] != 0.f) {
            r.t = fmaxf(r.t, fminf(t1, t2));
            r.tmax = fminf(r.tmax, fmaxf(t1, t2));
        }
    }
    return !(r.t > r.tmax);
}

// One step of the walk from r.t: an occupied cell's link, with its step dt
// (_intersect_aabb_unit, render_util.cuh:448-461, plus 1e-2), or -1 after
// skipping an empty block (compute_skip_dist_nn, :23-48).  Advances r.t.
template <typename DT>
__device__ __forceinline__ int nn_step(const GridView<DT>& g, NnRay& r, float& dt) {
    float p[3];
    int l[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        p[a] = r.o[a] + r.t * r.d[a];
        p[a] = fminf(fmaxf(p[a], 0.f), static_cast<float>(g.size[a]) - 1.f);
        l[a] = min(static_cast<int>(p[a]), g.size[a] - 1);
        p[a] -= static_cast<float>(l[a]);
    }
    const int link = g.links[l[0] * g.stride_x + l[1] * g.stride_y + l[2]];
    if (g.valid(link)) {
        float e = 1e9f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float t1 = -p[a] * r.inv[a];
            const float t2 = t1 + r.inv[a];
            e = fminf(e, fmaxf(t1, t2));
        }
        dt = e + 1e-2f;
        r.t += dt;
        return link;
    }
    const uint32_t dist = link < 0 ? static_cast<uint32_t>(-link) : 1u;
    const int shift = static_cast<int>(min(dist - 1u, 30u));
    const int side = 1 << shift;
    float smax = 1e9f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int ul = ((l[a] >> shift) << shift) - l[a];
        const float t1 = (static_cast<float>(ul) - p[a]) * r.inv[a];
        const float t2 = (static_cast<float>(ul + side) - p[a]) * r.inv[a];
        if (r.d[a] != 0.f) smax = fminf(smax, fmaxf(t1, t2));
    }
    r.t += fmaxf(smax, 0.f) + 1e-2f;
    return -1;
}

// Channel k's colour + 0.5 of a row, summed coefficient by coefficient.
template <typename DT, int B>
__device__ __forceinline__ float nn_colour(const GridView<DT>& g, int link, int k,
                                           const float* basis) {
    const int64_t base = static_cast<int64_t>(link) * g.sstride + k * B;
    float v = 0.5f;
#pragma unroll
    for (int i = 0; i < B; ++i) v += basis[i] * ld(g.sh, base + i);
    return v;
}

// svox1 forward (trace_ray, :50-160): rgb_out [n][3], one thread per ray.
template <typename DT, int B>
__global__ __launch_bounds__(kNnBlock) void k_render_svox1(GridView<DT> g, OptView o,
                                                           const float* __restrict__ origins,
                                                           const float* __restrict__ dirs,
                                                           int64_t n, float* __restrict__ rgb_out) {
    const int64_t ray = static_cast<int64_t>(blockIdx.x) * kNnBlock + threadIdx.x;
    if (ray >= n) return;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    float* out = rgb_out + 3 * ray;
    NnRay r;
    if (!nn_setup(g, wo, wd, false, r)) {
        out[0] = out[1] = out[2] = o.bg;
        return;
    }
    const RayCtx<B> rc(wd);
    float acc[3] = {0.f, 0.f, 0.f}, log_t = 0.f;
    int guard = 0;
    while (r.t < r.tmax && ++guard < kMaxIters) {
        float dt;
        const int link = nn_step(g, r, dt);
        if (link < 0) continue;
        float sigma = ld(g.density, static_cast<int64_t>(link) * g.dstride);
        if (o.last_opaque && r.t + o.step > r.tmax) sigma += 1e9f;
        if (!(sigma > o.sigma_thresh)) continue;
        const float lt = -(dt * r.ds) * sigma;
        const float tr = expf(log_t);
        const float w = tr * (1.f - expf(lt));
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += w * fmaxf(nn_colour<DT, B>(g, link, k, rc.basis), 0.f);
        log_t += lt;
        if (tr <= o.stop) {
            // Full opacity: the colour is renormalised and the background
            // left out (:137-145).
            const float scale = static_cast<float>(1.0 / (1.0 - static_cast<double>(tr)));
            out[0] = acc[0] * scale;
            out[1] = acc[1] * scale;
            out[2] = acc[2] * scale;
            return;
        }
    }
    const float tb = expf(log_t) * o.bg;
    out[0] = acc[0] + tb;
    out[1] = acc[1] + tb;
    out[2] = acc[2] + tb;
}

// svox1 backward (trace_ray_backward, :162-275), one ray per lane.
// grad_in: dL/drgb [n][3], or with rgb_residual the targets (the fused MSE
// step: g = (color_cache - grad_in) * norm).  It stops after the sample whose
// transmittance after it is <= stop_thresh, as the reference's does (its
// forward stops one counted sample later and renormalises; the backward
// differentiates neither).  The lanes step their walks together under a
// wave-uniform loop, and after each step the wave scatters that step's
// counted samples one by one with a lane per SH coefficient (and one for the
// density): each sample's 27 gradient atomics then fall in one row in one
// instruction, instead of 27 instructions whose lanes each hit a row of
// their own.
template <typename DT, int B>
__global__ __launch_bounds__(kNnBlock) void k_render_svox1_backward(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, int64_t n, const float* __restrict__ grad_in,
        const float* __restrict__ color_cache, int rgb_residual, float norm, GradOut go) {
    constexpr int C = 3 * B;
    const int64_t ray = static_cast<int64_t>(blockIdx.x) * kNnBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool active = ray < n;
    float wd[3] = {0.f, 0.f, 1.f};
    NnRay r;
    float gq[3] = {0.f, 0.f, 0.f}, accum = 0.f;
    if (active) {
        const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
#pragma unroll
        for (int a = 0; a < 3; ++a) wd[a] = dirs[3 * ray + a];
        active = nn_setup(g, wo, wd, true, r);
    }
    if (active) {
        const float* cc = color_cache + 3 * ray;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            gq[k] = rgb_residual ? (cc[k] - grad_in[3 * ray + k]) * norm : grad_in[3 * ray + k];
        accum = cc[0] * gq[0] + cc[1] * gq[1] + cc[2] * gq[2];
    }
    const RayCtx<B> rc(wd);
    float log_t = 0.f;
    int guard = 0;
    while (__ballot(active) != 0ull && ++guard < kMaxIters) {
        int link = -1;
        bool counted = false;
        float gw[3] = {0.f, 0.f, 0.f}, gdens = 0.f;
        if (active && !(r.t < r.tmax)) active = false;
        if (active) {
            float dt;
            link = nn_step(g, r, dt);
            if (link >= 0) {
                float sigma = ld(g.density, static_cast<int64_t>(link) * g.dstride);
                if (o.last_opaque && r.t + o.step > r.tmax) sigma += 1e9f;
                if (sigma > o.sigma_thresh) {
                    counted = true;
                    const float dtw = dt * r.ds;
                    const float lt = -dtw * sigma;
                    const float w = expf(log_t) * (1.f - expf(lt));
                    float total = 0.f;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float c = nn_colour<DT, B>(g, link, k, rc.basis);
                        if (c > 0.f) {
                            total += c * gq[k];
                            gw[k] = w * gq[k];
                        }
                    }
                    log_t += lt;
                    accum -= w * total;
                    gdens = dtw * (total * expf(log_t) - accum);
                    if (expf(log_t) <= o.stop) active = false;
                }
            }
        }
        unsigned long long todo = __ballot(counted);
        while (todo != 0ull) {
            const int j = __ffsll(todo) - 1;
            todo &= todo - 1ull;
            const int lj = __builtin_amdgcn_readlane(link, j);
            const float g0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gw[0]), j));
            const float g1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gw[1]), j));
            const float g2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gw[2]), j));
            const float gdj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gdens), j));
            float bl = 0.f;
#pragma unroll
            for (int i = 0; i < B; ++i) {
                const float bi =
                    __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rc.basis[i]), j));
                if (lane % B == i) bl = bi;
            }
            if (lane < C) {
                const int k = lane / B;
                const float gk = k == 0 ? g0 : (k == 1 ? g1 : g2);
                if (gk != 0.f && go.gs != nullptr)
                    atomic_add_f32(go.gs + static_cast<int64_t>(lj) * go.gss + lane, bl * gk);
            } else if (lane == C) {
                if (go.gd != nullptr) atomic_add_f32(go.gd + static_cast<int64_t>(lj) * go.gds, gdj);
                if (go.mask != nullptr) go.mask[lj] = 1;
            }
        }
    }
}

// t of lane k of a chunk whose lane 0 has t0: the float chain t0 + step + ...
// + step (k additions), by one fma when every addition is exact (see
// place_chunk), else replayed.
__device__ __forceinline__ float chain_t(float t0, int k, const OptView& o) {
    if (sum_chain_exact(t0, o.step, o.step_ulp_max))
        return fmaf(static_cast<float>(k), o.step, t0);
    float t = t0;
#pragma unroll
    for (int kk = 1; kk < 32; ++kk) {
        const float nx = t + o.step;
        t = (kk <= k) ? nx : t;
    }
    return t;
}

// Per-voxel maximum compositing weight of the pixel rays of one camera
// through a dense density volume (grid_trace_ray / grid_weight_render_kernel,
// misc_kernel.cu:184-334,418-445): every step of size step_size from the
// volume's entry (no empty-space skip), trilinear density, a sample with
// density > 1e-8 adds w = T (1 - exp(-ws sigma)) and raises the weight of the
// 8 corners of its cell to at least w; the ray stops after the first sample
// that leaves T below stop_thresh.  With last_sample_opaque the last sample
// before tmax gets density + 1e9 and T reset to 1 (weight 1).
//
// 32 lanes per pixel take 32 consecutive steps at a time; the log
// transmittance is replayed in step order with the sequential march's float
// subtractions, so the stop decision and every weight are those of a
// sequential walk.  Weights are non-negative floats, whose bit patterns order
// as integers: the per-corner maximum is an integer atomic max.
template <bool OPAQUE>
SVOX2_FWD_KERNEL void k_grid_weight_render(GridView<float> g, OptView o, svox2_camera_t cam,
                                           const float* __restrict__ data,
                                           int* __restrict__ weight_out) {
    const int64_t pix = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (pix >= static_cast<int64_t>(cam.width) * cam.height) return;
    const int iy = static_cast<int>(pix / cam.width), ix = static_cast<int>(pix % cam.width);
    float wd[3], wo[3];
    camera_ray(cam, ix, iy, wd, wo);
    Ray r;
    setup_ray(g, o, wo, wd, nullptr, r);
    if (r.tmin > r.tmax) return;
    const int sx = g.stride_x, sy = g.stride_y;
    float t0 = r.tmin, log_t = 0.f;
    int guard = 0;
    while (++guard < kMaxIters) {
        const float t = chain_t(t0, lane, o);
        const bool in_range = t <= r.tmax;
        float p[3];
        int l[3];
        float sigma = 0.f;
        int base = 0;
        if (in_range) {
            base = sample_cell(g, r, t, p, l);
            const float v[8] = {data[base], data[base + 1], data[base + sy], data[base + sy + 1],
                                data[base + sx], data[base + sx + 1], data[base + sx + sy],
                                data[base + sx + sy + 1]};
            sigma = trilerp8(v, p);
        }
        const bool last = OPAQUE && in_range && t + o.step > r.tmax;
        if (last) sigma += 1e9f;
        const bool contrib = in_range && sigma > 1e-8f;
        const float pcnt = contrib ? r.ws * sigma : 0.f;
        float lt = log_t;
        const int n_rep = wave_last_set(contrib);
        for (int j = 0; j < n_rep; ++j) {
            const float pj = group_lane(pcnt, j);
            if (j < lane) lt -= pj;
        }
        if (last) lt = 0.f;
        const float lt_after = lt - pcnt;
        const uint32_t stop_mask = group_ballot(contrib && lt_after < o.log_stop);
        const int first_stop = stop_mask ? (__ffs(stop_mask) - 1) : 32;
        if (contrib && lane <= first_stop) {
            const float w = __expf(lt) * (1.f - __expf(-pcnt));
            const int wb = __float_as_int(w);
            int* q = weight_out + base;
            atomicMax(q, wb);
            atomicMax(q + 1, wb);
            atomicMax(q + sy, wb);
            atomicMax(q + sy + 1, wb);
            atomicMax(q + sx, wb);
            atomicMax(q + sx + 1, wb);
            atomicMax(q + sx + sy, wb);
            atomicMax(q + sx + sy + 1, wb);
        }
        const uint32_t range_mask = group_ballot(in_range);
        if (first_stop < 32 || range_mask != 0xFFFFFFFFu) break;
        log_t = group_lane(lt_after, 31);
        t0 = group_lane(t, 31) + o.step;
    }
}


template <typename DT, int B>
SVOX2_FWD_KERNEL void k_render_image(
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
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, nullptr, r);
    float out[3];
    const float log_t = march_forward<DT, B, false>(g, o, r, rc, lane, out, nullptr);
    if (lane < 3) rgb_out[3 * ray + lane] = pick3(out, lane);
    if (log_t_out != nullptr && lane == 0) log_t_out[ray] = log_t;
}

// Pixels of many image requests per launch.  blockIdx.y selects the tile;
// tiles past *n_tiles, and pixels past a tile's n_pixels, exit at once.
// (One workgroup per item lets the hardware balance uneven pixels; a fixed
// grid striding over the items measured slower.)
template <typename DT, int B>
SVOX2_FWD_KERNEL void k_render_tiles(const svox2_grid_t* __restrict__ grids, int n_grids,
                                     const svox2_camera_t* __restrict__ cams, int n_cams,
                                     OptView o, const svox2_tile_t* __restrict__ tiles,
                                     const int32_t* __restrict__ n_tiles) {
    const int t = blockIdx.y;
    if (t >= *n_tiles) return;
    const svox2_tile_t tile = tiles[t];
    const int64_t local = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (local >= tile.n_pixels) return;
    if (tile.grid < 0 || tile.grid >= n_grids || tile.camera < 0 || tile.camera >= n_cams) return;
    const svox2_camera_t cam = cams[tile.camera];
    const int64_t pixel = tile.pixel0 + local;
    if (cam.width <= 0 || pixel >= static_cast<int64_t>(cam.width) * cam.height) return;
    const GridView<DT> g = make_view<DT>(grids[tile.grid]);
    const int ix = static_cast<int>(pixel % cam.width);
    const int iy = static_cast<int>(pixel / cam.width);
    float wd[3], wo[3];
    camera_ray(cam, ix, iy, wd, wo);
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, nullptr, r);
    float out[3];
    march_forward<DT, B, false>(g, o, r, rc, lane, out, nullptr);
    if (lane < 3) tile.rgb_out[3 * pixel + lane] = pick3(out, lane);
}

template <typename DT, int B>
SVOX2_BWD_KERNEL void k_render_backward(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        const float* __restrict__ grad_out, const float* __restrict__ color_cache,
        const uint32_t* __restrict__ order, GradOut go) {
    const int64_t slot = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (slot >= n) return;
    const int64_t ray = order != nullptr ? static_cast<int64_t>(order[slot]) : slot;
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    const float gq[3] = {grad_out[3 * ray], grad_out[3 * ray + 1], grad_out[3 * ray + 2]};
    const float* cc = color_cache + 3 * ray;
    const float accum = fmaf(cc[0], gq[0], fmaf(cc[1], gq[1], cc[2] * gq[2]));
    __shared__ BwdShared sh;
    BwdRay<B> st(rc, lane, gq, accum, 0.f, &sh, go);
    march_backward<DT, B>(g, o, r, rc, lane, st, go);
}

// Backward of the fused MSE step for the rays whose samples the forward
// recorded (state 1): the same gradient as k_render_backward_mse, from the
// records, in a kernel without the march code (fewer registers, more waves).
// The ray's colour is the sum of its chunk colours (k_rec_colours) in march
// order plus the background term, as the forward forms it; the group of item
// < n writes it to rgb_out.  A segment m >= 1 of a split ray starts from the
// colour composited before its first chunk.
template <typename DT, int B>
SVOX2_REPLAY_KERNEL void k_render_backward_replay(
        GridView<DT> g, const float* __restrict__ dirs, int64_t n,
        const float* __restrict__ rgb_gt, float* __restrict__ rgb_out,
        const float* __restrict__ log_t_fwd, float beta_per_ray, float sparsity, float norm,
        const uint32_t* __restrict__ order, GradOut go, RecStream rs) {
    const int64_t item = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    int64_t ray, pos;
    int max_chunks = kMaxIters;
    int seg_chunk = -1;  // recorded chunks before the segment (segments m >= 1)
    if (item < n) {  // a ray from its start; a split ray's first segment only
        ray = order != nullptr ? static_cast<int64_t>(order[item]) : item;
        if (rs.state[ray] != 1) return;  // re-marched by k_render_backward_mse
        pos = ray * rs.fixed;
        if (rs.split[ray] >= 0) max_chunks = kRecSegChunks;
    } else {  // a listed segment m >= 1 of a split ray
        const int64_t e = item - n;
        if (e >= rs.parts * rs.spart) return;
        const int64_t part = e / rs.spart;
        if (e - part * rs.spart >= min(static_cast<int64_t>(rs.next_block[kRecParts + part]),
                                       rs.spart))
            return;  // past the partition's entries
        const int2 sm = rs.seg_list[e];
        if (sm.x < 0) return;  // entries of a ray whose segments did not fit
        ray = sm.x;
        const int m = sm.y;
        pos = rs.seg[ray * (kRecSegs - 1) + m - 1];
        if (pos < 0) return;  // the ray ends before this segment
        if (m < kRecSegs - 1) max_chunks = kRecSegChunks;
        seg_chunk = m * kRecSegChunks;
    }
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const RayCtx<B> rc(wd);
    // The ray's colour, and the colour composited before the segment.
    const float4 info = rs.ray_info[ray];
    const int first = __float_as_int(info.x), cnt = __float_as_int(info.y);
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 v0 = lane < cnt ? rs.chunk_rgb[first + lane] : zero4;
    const float4 v1 = lane + kGroup < cnt ? rs.chunk_rgb[first + kGroup + lane] : zero4;
    float c3[3] = {0.f, 0.f, 0.f}, before[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < cnt; ++k) {
        if (k == seg_chunk) {
#pragma unroll
            for (int c = 0; c < 3; ++c) before[c] = c3[c];
        }
        const float4 v = k < kGroup ? v0 : v1;
        c3[0] += group_bcast(v.x, k & (kGroup - 1));
        c3[1] += group_bcast(v.y, k & (kGroup - 1));
        c3[2] += group_bcast(v.z, k & (kGroup - 1));
    }
    float gq[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        c3[c] += info.z;
        gq[c] = (c3[c] - rgb_gt[3 * ray + c]) * norm;
    }
    if (item < n && lane < 3) rgb_out[3 * ray + lane] = pick3(c3, lane);
    float accum = fmaf(c3[0], gq[0], fmaf(c3[1], gq[1], c3[2] * gq[2]));
    if (beta_per_ray > 0.f) {
        const float tin = __expf(log_t_fwd[ray]);
        accum += beta_per_ray * (1.f - tin / (1.f - tin + 1e-3f));
    }
    if (seg_chunk >= 0)  // accum after the samples before the segment
        accum -= fmaf(before[0], gq[0], fmaf(before[1], gq[1], before[2] * gq[2]));
    __shared__ BwdShared sh;
    BwdRay<B> st(rc, lane, gq, accum, sparsity, &sh, go);
    replay_backward<DT, B>(g, rs, pos, max_chunks, lane, st, go);
}

// Backward of the fused MSE step: dL/drgb = (rgb - gt) * 2 / (3n) from the
// forward colours (render_ray_backward_kernel with grad_out_is_rgb,
// render_lerp_kernel_cuvol.cu:735-741), plus the beta term from the forward's
// final log transmittance (:257-262) and the Cauchy sparsity term (:345-348).
template <typename DT, int B, int WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES)))
void k_render_backward_mse(
        GridView<DT> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        const float* __restrict__ rgb_gt, const float* __restrict__ rgb,
        const float* __restrict__ log_t_fwd, float beta_per_ray, float sparsity, float norm,
        const uint32_t* __restrict__ order, GradOut go, RecStream rs, ChunkLog log) {
    const int64_t slot = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kGroup;
    const int lane = threadIdx.x % kGroup;
    if (slot >= n) return;
    const int64_t ray = order != nullptr ? static_cast<int64_t>(order[slot]) : slot;
    if (rs.rec != nullptr && rs.state[ray] == 1) return;  // replayed by k_render_backward_replay
    const float wo[3] = {origins[3 * ray], origins[3 * ray + 1], origins[3 * ray + 2]};
    const float wd[3] = {dirs[3 * ray], dirs[3 * ray + 1], dirs[3 * ray + 2]};
    const RayCtx<B> rc(wd);
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * ray : nullptr, r);
    float c3[3], gq[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        c3[c] = rgb[3 * ray + c];
        gq[c] = (c3[c] - rgb_gt[3 * ray + c]) * norm;
    }
    float accum = fmaf(c3[0], gq[0], fmaf(c3[1], gq[1], c3[2] * gq[2]));
    if (beta_per_ray > 0.f) {
        const float tin = __expf(log_t_fwd[ray]);
        accum += beta_per_ray * (1.f - tin / (1.f - tin + 1e-3f));
    }
    __shared__ BwdShared sh;
    BwdRay<B> st(rc, lane, gq, accum, sparsity, &sh, go);
    march_backward<DT, B, (WAVES <= 3)>(g, o, r, rc, lane, st, go, log, ray);
}

// Counting sort of the rays by descending cost in two launches: a histogram,
// then a scatter whose workgroups each derive the bins' start offsets from the
// histogram.  The workspace holds two histograms (a call that sorts twice
// uses one each, so both can be cleared by its first kernel), the per-bin
// cursors the scatter reserves ranges with, then the per-ray costs and the
// resulting order (n words each).
constexpr int kSortBlock = 256;
constexpr int kOrderHeadWords = 3 * kCostBins;  // hist 0, hist 1, cursors

// LDS histogram per workgroup, added into hist; workgroup 0 also clears the
// cursors the scatter will use.
__global__ __launch_bounds__(kSortBlock) void k_cost_hist(const uint32_t* __restrict__ cost,
                                                          int64_t n, uint32_t* __restrict__ hist,
                                                          uint32_t* __restrict__ cursor) {
    __shared__ uint32_t h[kCostBins];
    for (int b = threadIdx.x; b < kCostBins; b += kSortBlock) h[b] = 0u;
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b < kCostBins; b += kSortBlock) cursor[b] = 0u;
    __syncthreads();
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kSortBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kSortBlock)
        atomicAdd(&h[cost[i]], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kCostBins; b += kSortBlock)
        if (h[b] != 0u) atomicAdd(&hist[b], h[b]);
}

// Each workgroup places kScatterRays consecutive rays.  It computes where
// every bin starts (the number of rays in higher bins: an exclusive scan of
// the histogram over descending bins, in LDS), counts its rays per bin,
// reserves its range of every bin with one atomic on that bin's cursor, then
// ranks them.  Few workgroups per batch keep the reservations on a popular
// bin (rays that miss the grid share bin 0) from queueing on one address.
constexpr int64_t kScatterRays = 4096;

__global__ __launch_bounds__(kSortBlock) void k_cost_scatter(const uint32_t* __restrict__ cost,
                                                             int64_t n,
                                                             const uint32_t* __restrict__ hist,
                                                             uint32_t* __restrict__ cursor,
                                                             uint32_t* __restrict__ order) {
    constexpr int kPer = kCostBins / kSortBlock;
    __shared__ uint32_t start[kCostBins];
    __shared__ uint32_t part[kSortBlock];
    __shared__ uint32_t h[kCostBins];
    const int t = threadIdx.x;
    // Thread t holds descending positions kPer * t .. kPer * t + kPer - 1.
    uint32_t c[kPer], sum = 0u;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        c[j] = hist[kCostBins - 1 - (kPer * t + j)];
        sum += c[j];
    }
    part[t] = sum;
    for (int b = t; b < kCostBins; b += kSortBlock) h[b] = 0u;
    __syncthreads();
    for (int off = 1; off < kSortBlock; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t at = part[t] - sum;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        start[kCostBins - 1 - (kPer * t + j)] = at;
        at += c[j];
    }
    const int64_t begin = static_cast<int64_t>(blockIdx.x) * kScatterRays;
    const int64_t end = min(n, begin + kScatterRays);
    for (int64_t i = begin + t; i < end; i += kSortBlock) atomicAdd(&h[cost[i]], 1u);
    __syncthreads();
    for (int b = t; b < kCostBins; b += kSortBlock)
        if (h[b] != 0u) h[b] = start[b] + atomicAdd(&cursor[b], h[b]);
    __syncthreads();
    for (int64_t i = begin + t; i < end; i += kSortBlock)
        order[atomicAdd(&h[cost[i]], 1u)] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(kSortBlock) void k_zero_words(uint32_t* __restrict__ p, int n) {
    const int i = blockIdx.x * kSortBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Sort key of the forward: the length of the ray's [tmin, tmax] segment in
// voxels, scaled to the cost bins (bins_per_voxel = bins / grid diagonal).
// The forward's work per ray is not known before it runs; this length ranks
// rays close to their measured chunk counts (empty space is skipped in a few
// chunks, but a longer segment crosses more of it and more of the occupied
// shell), and ordering by it pairs rays of similar length in a wave and
// starts the long ones first.
__global__ __launch_bounds__(kSortBlock) void k_ray_length_key(
        GridView<float> g, OptView o, const float* __restrict__ origins,
        const float* __restrict__ dirs, const float* __restrict__ clip, int64_t n,
        float bins_per_voxel, uint32_t* __restrict__ key, uint32_t* __restrict__ clear) {
    // Workgroup 0 clears both histograms of the ordering workspace.
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b < 2 * kCostBins; b += kSortBlock) clear[b] = 0u;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kSortBlock + threadIdx.x;
    if (i >= n) return;
    const float wo[3] = {origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]};
    const float wd[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    Ray r;
    setup_ray(g, o, wo, wd, clip ? clip + 2 * i : nullptr, r);
    const float len = r.tmax - r.tmin;
    uint32_t b = 0;
    if (len > 0.f)  // NaN-safe: a NaN length keeps bin 0
        b = static_cast<uint32_t>(fminf(len * bins_per_voxel, static_cast<float>(kCostBins - 1)));
    key[i] = b;
}

struct OrderWorkspace {
    uint32_t* hist = nullptr;    // two histograms of kCostBins words
    uint32_t* cursor = nullptr;  // kCostBins words
    uint32_t* cost = nullptr;
    uint32_t* order = nullptr;
};

size_t order_workspace_bytes(int64_t n) {
    const size_t words = static_cast<size_t>(kOrderHeadWords) + 2 * static_cast<size_t>(n);
    return (words * 4 + 255) / 256 * 256;
}

OrderWorkspace order_workspace(void* ws, int64_t n) {
    OrderWorkspace w;
    uint32_t* p = static_cast<uint32_t*>(ws);
    w.hist = p;
    w.cursor = p + 2 * kCostBins;
    w.cost = p + kOrderHeadWords;
    w.order = w.cost + n;
    return w;
}

// Fused-call workspace: the ordering workspace, then the record stream:
// per-ray state (n words), the overflow counter, and the record slots (n *
// (kRecFixed + kRecOverflow) slots of 96 bytes).
constexpr int64_t kRecFixed = 8;      // slots per ray in the fixed region
constexpr int64_t kRecOverflow = 40;  // overflow slots per ray
constexpr int64_t kRecSlotBytes = 96;

inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// Segment list entries: three extra segments per ray of the batch (rays
// whose segments do not fit are replayed whole).
int64_t seg_cap_of(int64_t n) { return 3 * n; }

// [order | state (n) | seg (n (kRecSegs - 1)) | split (n) | seg_list | counters |
//  ray_info (n) | clist (kDeferCap n) | chunk_rgb (kDeferCap n) | records]
size_t fused_workspace_bytes(int64_t n) {
    const int64_t order = static_cast<int64_t>(order_workspace_bytes(n));
    const int64_t state = align256(4 * n);
    const int64_t seg = align256(4 * n * (kRecSegs - 1));
    const int64_t split = align256(4 * n);
    const int64_t list = align256(8 * seg_cap_of(n));
    const int64_t counter = align256(4 * 3 * kRecParts);
    const int64_t info = align256(16 * n);
    const int64_t clist = align256(8 * kDeferCap * n);
    const int64_t crgb = align256(16 * kDeferCap * n);
    const int64_t recs = n * (kRecFixed + kRecOverflow) * kRecSlotBytes;
    return static_cast<size_t>(order + state + seg + split + list + counter + info + clist +
                               crgb + recs);
}

// Fused-call workspace with the chunk log (batches that re-march): the
// ordering workspace, then the log records (n * kLogChunks * kLogWords
// words) and the per-ray heads (n int4).
size_t logged_workspace_bytes(int64_t n) {
    const int64_t order = static_cast<int64_t>(order_workspace_bytes(n));
    const int64_t recs = align256(n * kLogChunks * kLogWords * 4);
    const int64_t heads = align256(16 * n);
    return static_cast<size_t>(order + recs + heads);
}

ChunkLog chunk_log(void* ws, int64_t n) {
    char* p = static_cast<char*>(ws) + order_workspace_bytes(n);
    ChunkLog log;
    log.rec = reinterpret_cast<float*>(p);
    log.head = reinterpret_cast<int4*>(p + align256(n * kLogChunks * kLogWords * 4));
    return log;
}

RecStream record_stream(void* ws, int64_t n) {
    char* p = static_cast<char*>(ws) + order_workspace_bytes(n);
    RecStream rs{};
    rs.state = reinterpret_cast<int32_t*>(p);
    p += align256(4 * n);
    rs.seg = reinterpret_cast<int32_t*>(p);
    p += align256(4 * n * (kRecSegs - 1));
    rs.split = reinterpret_cast<int32_t*>(p);
    p += align256(4 * n);
    rs.seg_list = reinterpret_cast<int2*>(p);
    rs.seg_cap = seg_cap_of(n);
    p += align256(8 * seg_cap_of(n));
    rs.next_block = reinterpret_cast<uint32_t*>(p);
    p += align256(4 * 3 * kRecParts);
    rs.ray_info = reinterpret_cast<float4*>(p);
    p += align256(16 * n);
    rs.clist = reinterpret_cast<int2*>(p);
    rs.clist_cap = kDeferCap * n;
    p += align256(8 * kDeferCap * n);
    rs.chunk_rgb = reinterpret_cast<float4*>(p);
    p += align256(16 * kDeferCap * n);
    rs.rec = reinterpret_cast<float4*>(p);
    rs.fixed = kRecFixed;
    rs.overflow_base = n * kRecFixed;
    rs.overflow_slots = n * kRecOverflow;
    // Counter partitions: one per 1024 rays of batches of 16384 rays or more,
    // at most kRecParts, so that each partition's share of the overflow
    // region and lists evens out over its rays; smaller batches share one
    // counter each (their atomics are few, and their overflow region is too
    // small to split).
    rs.parts = n < 16384 ? 1 : static_cast<int>(std::min<int64_t>(n / 1024, kRecParts));
    rs.opart = rs.overflow_slots / rs.parts;
    rs.spart = rs.seg_cap / rs.parts;
    rs.cpart = rs.clist_cap / rs.parts;
    return rs;
}

// Gradient descriptor checks of the backward entry points: the scatter
// addresses a row as link * stride with a 32-bit byte stride.
int check_grads(const svox2_grads_t* grads, int sh_dim, GradOut& go) {
    SVOX2_REQUIRE(grads != nullptr, "grads is NULL");
    SVOX2_REQUIRE(grads->grad_density_stride >= 0 && grads->grad_sh_stride >= 0,
                  "negative gradient stride");
    go = GradOut{grads->grad_density_out, grads->grad_sh_out, grads->mask_out,
                 grad_dstride(*grads), grad_sstride(*grads, sh_dim)};
    SVOX2_REQUIRE(go.gss >= sh_dim, "grad_sh_stride below sh_data_dim");
    SVOX2_REQUIRE(go.gds < (int64_t{1} << 30) && go.gss < (int64_t{1} << 30),
                  "gradient row strides must be below 2^30 elements");
    return ok();
}

int check_order_workspace(const void* workspace, size_t bytes, int64_t n) {
    if (workspace == nullptr) return ok();
    SVOX2_REQUIRE(bytes >= order_workspace_bytes(n),
                  "workspace of %zu bytes is below the %zu bytes %lld rays need", bytes,
                  order_workspace_bytes(n), (long long)n);
    SVOX2_REQUIRE(n <= static_cast<int64_t>(UINT32_MAX),
                  "ray ordering supports at most 2^32 - 1 rays per call");
    return ok();
}

// Batches at least this large are ordered: four times the rays the current
// device holds resident (two per wave, at the forward's occupancy).
int64_t order_min_rays() {
    static int64_t cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return INT64_MAX;
    if (cached[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            return INT64_MAX;
        cached[dev] = 4 * static_cast<int64_t>(cus) * 4 * SVOX2_FWD_WAVES * (64 / kGroup);
    }
    return cached[dev];
}

// Sorts the rays by the costs the forward wrote (w.cost) into w.order.
// Sorts rays by descending cost into w.order with histogram `which` (0 or 1),
// which must be clear when `cleared` says so (else it is cleared here).
void launch_cost_order(const uint32_t* cost, const OrderWorkspace& w, int64_t n, hipStream_t s,
                       int which = 0, bool cleared = false) {
    uint32_t* hist = w.hist + which * kCostBins;
    if (!cleared)
        hipLaunchKernelGGL(k_zero_words, dim3(blocks_for(kCostBins, kSortBlock)), dim3(kSortBlock), 0,
                           s, hist, kCostBins);
    const unsigned blocks = static_cast<unsigned>((n + kScatterRays - 1) / kScatterRays);
    hipLaunchKernelGGL(k_cost_hist, dim3(blocks), dim3(kSortBlock), 0, s, cost, n, hist, w.cursor);
    hipLaunchKernelGGL(k_cost_scatter, dim3(blocks), dim3(kSortBlock), 0, s, cost, n, hist,
                       w.cursor, w.order);
}

// Orders the rays of a forward by descending segment length (keys into
// w.cost, then the counting sort into w.order).
void launch_length_order(const svox2_grid_t* grid, const OptView& o, const float* origins,
                         const float* dirs, const float* clip, int64_t n,
                         const OrderWorkspace& w, hipStream_t s) {
    const double diag = sqrt(static_cast<double>(grid->size[0]) * grid->size[0] +
                             static_cast<double>(grid->size[1]) * grid->size[1] +
                             static_cast<double>(grid->size[2]) * grid->size[2]);
    hipLaunchKernelGGL(k_ray_length_key, dim3(blocks_for(n, kSortBlock)), dim3(kSortBlock), 0, s,
                       make_view<float>(*grid), o, origins, dirs, clip, n,
                       static_cast<float>((kCostBins - 1) / diag), w.cost, w.hist);
    launch_cost_order(w.cost, w, n, s, 0, true);
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
    {
        // step = m * 2^(e_step - 150) with a 24-bit m: a whole multiple of
        // 2^(e - 150) exactly when e <= e_step + (trailing zeros of m).
        uint32_t bits;
        memcpy(&bits, &o.step, 4);
        const uint32_t m = (bits & 0x7FFFFFu) | 0x800000u;
        o.step_ulp_max = static_cast<int>((bits >> 23) & 0xFFu) + __builtin_ctz(m);
    }
    o.sigma_thresh = opt->sigma_thresh;
    o.log_stop = log_stop_threshold(opt->stop_thresh);
    o.stop = opt->stop_thresh;
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

// Forward render of a ray batch; with an ordering workspace, batches at or
// above the ordering threshold run in descending order of segment length.
int render_rays(const svox2_grid_t* grid, const float* origins, const float* dirs,
                const float* ray_clip, int64_t n_rays, const svox2_render_options_t* opt,
                float* rgb_out, float* log_transmit_out, uint32_t* ray_cost_out,
                void* workspace, size_t workspace_bytes, void* stream, const char* what) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_out != nullptr, "rgb_out is NULL");
    if ((st = check_order_workspace(workspace, workspace_bytes, n_rays))) return st;
    // A background is composited behind the grid's own colours, which are
    // rendered with brightness 0 (render_lerp_kernel_cuvol.cu:43,111-113).
    const bool background = grid_has_background(grid);
    if (background) {
        if ((st = check_background(grid, opt))) return st;
        SVOX2_REQUIRE(log_transmit_out != nullptr,
                      "a grid with a background needs log_transmit_out (n floats)");
        o.bg = 0.f;
    }
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    const hipStream_t s = as_stream(stream);
    OrderWorkspace w;
    if (workspace != nullptr && n_rays >= order_min_rays()) {
        w = order_workspace(workspace, n_rays);
        launch_length_order(grid, o, origins, dirs, ray_clip, n_rays, w, s);
    }
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_rays<DT, B, false>), dim3(blocks), dim3(kBlock), 0, s, g, o,
                           origins, dirs, ray_clip, n_rays, rgb_out, log_transmit_out,
                           (int32_t*)nullptr, ray_cost_out, RecStream{},
                           static_cast<const uint32_t*>(w.order), ChunkLog{});
    });
    if (st) return st;
    if (background && (st = launch_background(grid, opt, origins, dirs, nullptr, n_rays,
                                              log_transmit_out, rgb_out, nullptr, s)))
        return st;
    return check_launch(what);
}

extern "C" int svox2_volume_render_cuvol(const svox2_grid_t* grid, const float* origins,
                                         const float* dirs, const float* ray_clip,
                                         int64_t n_rays, const svox2_render_options_t* opt,
                                         float* rgb_out, float* log_transmit_out,
                                         uint32_t* ray_cost_out, void* stream) {
    return render_rays(grid, origins, dirs, ray_clip, n_rays, opt, rgb_out, log_transmit_out,
                       ray_cost_out, nullptr, 0, stream, "volume_render_cuvol");
}

extern "C" int svox2_volume_render_cuvol_ordered(const svox2_grid_t* grid, const float* origins,
                                                 const float* dirs, const float* ray_clip,
                                                 int64_t n_rays,
                                                 const svox2_render_options_t* opt,
                                                 float* rgb_out, float* log_transmit_out,
                                                 uint32_t* ray_cost_out, void* workspace,
                                                 size_t workspace_bytes, void* stream) {
    return render_rays(grid, origins, dirs, ray_clip, n_rays, opt, rgb_out, log_transmit_out,
                       ray_cost_out, workspace, workspace_bytes, stream,
                       "volume_render_cuvol_ordered");
}

// Depth of a ray batch, mode kDepthExpected or kDepthSigma (threshold
// sigma_thresh); the render options' last_sample_opaque and background do
// not apply, as in the reference's depth kernels.
int render_depth(const svox2_grid_t* grid, const float* origins, const float* dirs,
                 const float* ray_clip, int64_t n_rays, const svox2_render_options_t* opt,
                 int mode, float sigma_thresh, float* depth_out, void* stream) {
    int st = check_grid(grid, false);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(depth_out != nullptr, "depth_out is NULL");
    o.last_opaque = 0;
    if (mode == kDepthSigma) {
        o.sigma_thresh = sigma_thresh;
        o.log_stop = -INFINITY;
    }
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    const hipStream_t s = as_stream(stream);
    auto launch = [&](auto g, auto dtp) {
        using DT = std::remove_pointer_t<decltype(dtp)>;
        if (mode == kDepthSigma)
            hipLaunchKernelGGL((k_render_depth<DT, kDepthSigma>), dim3(blocks), dim3(kBlock), 0,
                               s, g, o, origins, dirs, ray_clip, n_rays, depth_out);
        else
            hipLaunchKernelGGL((k_render_depth<DT, kDepthExpected>), dim3(blocks), dim3(kBlock),
                               0, s, g, o, origins, dirs, ray_clip, n_rays, depth_out);
    };
    if (grid->storage == SVOX2_STORAGE_BF16)
        launch(make_view<uint16_t>(*grid), (uint16_t*)nullptr);
    else
        launch(make_view<float>(*grid), (float*)nullptr);
    return check_launch(mode == kDepthSigma ? "volume_render_sigma_thresh"
                                            : "volume_render_expected_term");
}

extern "C" int svox2_grid_weight_render(const float* data, const int32_t size[3],
                                        const svox2_camera_t* cam, float step_size,
                                        float stop_thresh, int32_t last_sample_opaque,
                                        const float offset[3], const float scaling[3],
                                        float* grid_weight_out, void* stream) {
    SVOX2_REQUIRE(data != nullptr && size != nullptr && cam != nullptr && offset != nullptr &&
                      scaling != nullptr && grid_weight_out != nullptr,
                  "grid_weight_render: NULL argument");
    for (int i = 0; i < 3; ++i)
        SVOX2_REQUIRE(size[i] >= 2, "volume size[%d] = %d, must be >= 2", i, size[i]);
    const int64_t cells = static_cast<int64_t>(size[0]) * size[1] * size[2];
    SVOX2_REQUIRE(cells < INT32_MAX, "volume has too many cells");
    SVOX2_REQUIRE(cam->width >= 0 && cam->height >= 0, "negative image size");
    svox2_render_options_t ro{};
    ro.step_size = step_size;
    ro.stop_thresh = stop_thresh;
    OptView o;
    int st = make_opt(&ro, o);
    if (st) return st;
    const int64_t n_pix = static_cast<int64_t>(cam->width) * cam->height;
    if (n_pix == 0) return ok();
    GridView<float> g{};
    for (int i = 0; i < 3; ++i) {
        g.size[i] = size[i];
        g.offset[i] = offset[i];
        g.scaling[i] = scaling[i];
    }
    g.stride_x = size[1] * size[2];
    g.stride_y = size[2];
    const unsigned blocks = blocks_for(n_pix * kGroup, kBlock);
    int* out = reinterpret_cast<int*>(grid_weight_out);
    if (last_sample_opaque)
        hipLaunchKernelGGL((k_grid_weight_render<true>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, *cam, data, out);
    else
        hipLaunchKernelGGL((k_grid_weight_render<false>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, *cam, data, out);
    return check_launch("grid_weight_render");
}

extern "C" int svox2_volume_render_expected_term(const svox2_grid_t* grid, const float* origins,
                                                 const float* dirs, const float* ray_clip,
                                                 int64_t n_rays,
                                                 const svox2_render_options_t* opt,
                                                 float* depth_out, void* stream) {
    return render_depth(grid, origins, dirs, ray_clip, n_rays, opt, kDepthExpected, 0.f,
                        depth_out, stream);
}

extern "C" int svox2_volume_render_sigma_thresh(const svox2_grid_t* grid, const float* origins,
                                                const float* dirs, const float* ray_clip,
                                                int64_t n_rays,
                                                const svox2_render_options_t* opt,
                                                float sigma_thresh, float* depth_out,
                                                void* stream) {
    return render_depth(grid, origins, dirs, ray_clip, n_rays, opt, kDepthSigma, sigma_thresh,
                        depth_out, stream);
}

// nvol backend entry points (render_lerp_kernel_nvol.cu:411-502): the
// colour march's placement without its transmittance stop or opaque last
// sample.
namespace svox2amd {
namespace {

int nvol_opt(const svox2_grid_t* grid, const float* origins, const float* dirs, int64_t n_rays,
             const svox2_render_options_t* opt, OptView& o) {
    int st = check_grid(grid, true);
    if (st) return st;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    o.log_stop = -INFINITY;
    o.last_opaque = 0;
    return ok();
}

int nvol_backward(const svox2_grid_t* grid, const OptView& o, const float* origins,
                  const float* dirs, const float* ray_clip, int64_t n_rays, const float* grad_in,
                  const float* color_cache, int rgb_residual, float norm, float sparsity,
                  const GradOut& go, hipStream_t s) {
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    return dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_nvol_backward<DT, B>), dim3(blocks), dim3(kBlock), 0, s, g,
                           o, origins, dirs, ray_clip, n_rays, grad_in, color_cache,
                           rgb_residual, norm, sparsity, go);
    });
}

}  // namespace
}  // namespace svox2amd

extern "C" int svox2_volume_render_nvol(const svox2_grid_t* grid, const float* origins,
                                        const float* dirs, const float* ray_clip,
                                        int64_t n_rays, const svox2_render_options_t* opt,
                                        float* rgb_out, void* stream) {
    OptView o;
    int st = nvol_opt(grid, origins, dirs, n_rays, opt, o);
    if (st) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_out != nullptr, "rgb_out is NULL");
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    const hipStream_t s = as_stream(stream);
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_nvol<DT, B>), dim3(blocks), dim3(kBlock), 0, s, g, o,
                           origins, dirs, ray_clip, n_rays, rgb_out);
    });
    if (st) return st;
    return check_launch("volume_render_nvol");
}

extern "C" int svox2_volume_render_nvol_backward(const svox2_grid_t* grid, const float* origins,
                                                 const float* dirs, const float* ray_clip,
                                                 int64_t n_rays,
                                                 const svox2_render_options_t* opt,
                                                 const float* grad_out,
                                                 const float* color_cache,
                                                 const svox2_grads_t* grads, void* stream) {
    OptView o;
    int st = nvol_opt(grid, origins, dirs, n_rays, opt, o);
    if (st) return st;
    GradOut go;
    if ((st = check_grads(grads, grid->sh_data_dim, go))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(grad_out != nullptr, "grad_out is NULL");
    (void)color_cache;  // the reference's backward does not read it either
    if ((st = nvol_backward(grid, o, origins, dirs, ray_clip, n_rays, grad_out, nullptr, 0, 1.f,
                            0.f, go, as_stream(stream))))
        return st;
    return check_launch("volume_render_nvol_backward");
}

extern "C" int svox2_volume_render_nvol_fused(const svox2_grid_t* grid, const float* origins,
                                              const float* dirs, const float* ray_clip,
                                              int64_t n_rays,
                                              const svox2_render_options_t* opt,
                                              const float* rgb_gt, float beta_loss,
                                              float sparsity_loss, float* rgb_out,
                                              const svox2_grads_t* grads, void* stream) {
    (void)beta_loss;  // not supported by the reference's nvol (render_lerp_kernel_nvol.cu:466)
    OptView o;
    int st = nvol_opt(grid, origins, dirs, n_rays, opt, o);
    if (st) return st;
    GradOut go;
    if ((st = check_grads(grads, grid->sh_data_dim, go))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_gt != nullptr && rgb_out != nullptr, "rgb_gt / rgb_out are NULL");
    if ((st = svox2_volume_render_nvol(grid, origins, dirs, ray_clip, n_rays, opt, rgb_out, stream)))
        return st;
    const float norm = 2.f / static_cast<float>(3 * n_rays);
    if ((st = nvol_backward(grid, o, origins, dirs, ray_clip, n_rays, rgb_gt, rgb_out, 1, norm,
                            sparsity_loss, go, as_stream(stream))))
        return st;
    return check_launch("volume_render_nvol_fused");
}

// svox1 backend entry points (render_svox1_kernel.cu:334-419).  The reference
// takes no ray clips and no background for this backend: a clip is refused,
// a background ignored.
namespace svox2amd {
namespace {

int svox1_opt(const svox2_grid_t* grid, const float* origins, const float* dirs,
              const float* ray_clip, int64_t n_rays, const svox2_render_options_t* opt,
              OptView& o) {
    int st = check_grid(grid, true);
    if (st) return st;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    SVOX2_REQUIRE(ray_clip == nullptr, "the svox1 backend takes no ray clips");
    return ok();
}

int svox1_backward(const svox2_grid_t* grid, const OptView& o, const float* origins,
                   const float* dirs, int64_t n_rays, const float* grad_in,
                   const float* color_cache, int rgb_residual, float norm, const GradOut& go,
                   hipStream_t s) {
    const unsigned blocks = blocks_for(n_rays, kNnBlock);
    return dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_svox1_backward<DT, B>), dim3(blocks), dim3(kNnBlock), 0, s,
                           g, o, origins, dirs, n_rays, grad_in, color_cache, rgb_residual, norm,
                           go);
    });
}

}  // namespace
}  // namespace svox2amd

extern "C" int svox2_volume_render_svox1(const svox2_grid_t* grid, const float* origins,
                                         const float* dirs, const float* ray_clip,
                                         int64_t n_rays, const svox2_render_options_t* opt,
                                         float* rgb_out, void* stream) {
    OptView o;
    int st = svox1_opt(grid, origins, dirs, ray_clip, n_rays, opt, o);
    if (st) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_out != nullptr, "rgb_out is NULL");
    const unsigned blocks = blocks_for(n_rays, kNnBlock);
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_svox1<DT, B>), dim3(blocks), dim3(kNnBlock), 0,
                           as_stream(stream), g, o, origins, dirs, n_rays, rgb_out);
    });
    if (st) return st;
    return check_launch("volume_render_svox1");
}

extern "C" int svox2_volume_render_svox1_backward(const svox2_grid_t* grid, const float* origins,
                                                  const float* dirs, const float* ray_clip,
                                                  int64_t n_rays,
                                                  const svox2_render_options_t* opt,
                                                  const float* grad_out,
                                                  const float* color_cache,
                                                  const svox2_grads_t* grads, void* stream) {
    OptView o;
    int st = svox1_opt(grid, origins, dirs, ray_clip, n_rays, opt, o);
    if (st) return st;
    GradOut go;
    if ((st = check_grads(grads, grid->sh_data_dim, go))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(grad_out != nullptr && color_cache != nullptr,
                  "grad_out / color_cache are NULL");
    if ((st = svox1_backward(grid, o, origins, dirs, n_rays, grad_out, color_cache, 0, 1.f, go,
                             as_stream(stream))))
        return st;
    return check_launch("volume_render_svox1_backward");
}

extern "C" int svox2_volume_render_svox1_fused(const svox2_grid_t* grid, const float* origins,
                                               const float* dirs, const float* ray_clip,
                                               int64_t n_rays,
                                               const svox2_render_options_t* opt,
                                               const float* rgb_gt, float beta_loss,
                                               float sparsity_loss, float* rgb_out,
                                               const svox2_grads_t* grads, void* stream) {
    (void)beta_loss;  // not supported by the reference's svox1 (:395-396)
    (void)sparsity_loss;
    OptView o;
    int st = svox1_opt(grid, origins, dirs, ray_clip, n_rays, opt, o);
    if (st) return st;
    GradOut go;
    if ((st = check_grads(grads, grid->sh_data_dim, go))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_gt != nullptr && rgb_out != nullptr, "rgb_gt / rgb_out are NULL");
    if ((st = svox2_volume_render_svox1(grid, origins, dirs, nullptr, n_rays, opt, rgb_out,
                                        stream)))
        return st;
    const float norm = 2.f / static_cast<float>(3 * n_rays);
    if ((st = svox1_backward(grid, o, origins, dirs, n_rays, rgb_gt, rgb_out, 1, norm, go,
                             as_stream(stream))))
        return st;
    return check_launch("volume_render_svox1_fused");
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
                           (float*)nullptr, (float*)nullptr, stats_out, (uint32_t*)nullptr,
                           RecStream{}, (const uint32_t*)nullptr, ChunkLog{});
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
    const bool background = grid_has_background(grid);
    if (background) {
        if ((st = check_background(grid, opt))) return st;
        SVOX2_REQUIRE(log_transmit_out != nullptr,
                      "a grid with a background needs log_transmit_out (width * height floats)");
        o.bg = 0.f;
    }
    const unsigned blocks = blocks_for(n * kGroup, kBlock);
    const svox2_camera_t c = *cam;
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_image<DT, B>), dim3(blocks), dim3(kBlock), 0,
                           as_stream(stream), g, o, c, rgb_out, log_transmit_out);
    });
    if (st) return st;
    if (background && (st = launch_background(grid, opt, nullptr, nullptr, &c, n,
                                              log_transmit_out, rgb_out, nullptr,
                                              as_stream(stream))))
        return st;
    return check_launch("volume_render_cuvol_image");
}

extern "C" int svox2_volume_render_cuvol_tiles(const svox2_grid_t* grids, int32_t n_grids,
                                               int32_t storage, int32_t basis_dim,
                                               const svox2_camera_t* cameras, int32_t n_cameras,
                                               const svox2_render_options_t* opt,
                                               const svox2_tile_t* tiles, const int32_t* n_tiles,
                                               int32_t max_tiles, int32_t tile_pixels,
                                               void* stream) {
    OptView o;
    int st = make_opt(opt, o);
    if (st) return st;
    SVOX2_REQUIRE(grids != nullptr && cameras != nullptr && tiles != nullptr && n_tiles != nullptr,
                  "grids / cameras / tiles / n_tiles are NULL");
    SVOX2_REQUIRE(n_grids >= 1 && n_cameras >= 1, "no grids or cameras");
    SVOX2_REQUIRE(max_tiles >= 0 && max_tiles <= 65535, "max_tiles %d out of range", max_tiles);
    SVOX2_REQUIRE(tile_pixels >= 1 && tile_pixels <= (1 << 20), "tile_pixels %d out of range",
                  tile_pixels);
    SVOX2_REQUIRE(storage == SVOX2_STORAGE_F32 || storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", storage);
    if (max_tiles == 0) return ok();
    const dim3 grid_dim(blocks_for(static_cast<int64_t>(tile_pixels) * kGroup, kBlock),
                        static_cast<unsigned>(max_tiles));
    const hipStream_t s = as_stream(stream);
    auto launch = [&](auto dtp, auto b) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        hipLaunchKernelGGL((k_render_tiles<DT, B>), grid_dim, dim3(kBlock), 0, s, grids, n_grids,
                           cameras, n_cameras, o, tiles, n_tiles);
    };
#define SVOX2_TILE_CASE(B)                                                         \
    if (basis_dim == B) {                                                          \
        if (storage == SVOX2_STORAGE_BF16) launch((uint16_t*)nullptr, std::integral_constant<int, B>()); \
        else launch((float*)nullptr, std::integral_constant<int, B>());           \
        return check_launch("volume_render_cuvol_tiles");                        \
    }
    SVOX2_TILE_CASE(1)
    SVOX2_TILE_CASE(4)
    SVOX2_TILE_CASE(9)
#undef SVOX2_TILE_CASE
    return fail(SVOX2_ERR_UNSUPPORTED, "basis_dim %d not supported (1, 4 or 9)", basis_dim);
}

// A background's per-call buffers (16 bytes per ray) follow the ordering
// workspace.
size_t background_offset(int64_t n) { return (order_workspace_bytes(n) + 255) & ~size_t{255}; }
float* background_buffer(void* workspace, int64_t n) {
    return reinterpret_cast<float*>(static_cast<char*>(workspace) + background_offset(n));
}

extern "C" int svox2_volume_render_cuvol_backward(const svox2_grid_t* grid,
                                                  const float* origins, const float* dirs,
                                                  const float* ray_clip, int64_t n_rays,
                                                  const svox2_render_options_t* opt,
                                                  const float* grad_out,
                                                  const float* color_cache,
                                                  const svox2_grads_t* grads,
                                                  const uint32_t* ray_cost, void* workspace,
                                                  size_t workspace_bytes, void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    GradOut go;
    if ((st = check_grads(grads, grid->sh_data_dim, go))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(grad_out != nullptr && color_cache != nullptr,
                  "grad_out / color_cache are NULL");
    if ((st = check_order_workspace(workspace, workspace_bytes, n_rays))) return st;
    const bool background = grid_has_background(grid);
    float* bg_buf = nullptr;  // the grid's own colours [n][3], then its log transmittance [n]
    if (background) {
        if ((st = check_background(grid, opt))) return st;
        SVOX2_REQUIRE(workspace != nullptr &&
                          workspace_bytes >= svox2_background_workspace_bytes(n_rays),
                      "a grid with a background needs a workspace of "
                      "svox2_background_workspace_bytes(n_rays) = %zu bytes",
                      svox2_background_workspace_bytes(n_rays));
        bg_buf = background_buffer(workspace, n_rays);
    }
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    const hipStream_t s = as_stream(stream);
    OrderWorkspace w;
    if (ray_cost != nullptr && workspace != nullptr && n_rays >= order_min_rays()) {
        w = order_workspace(workspace, n_rays);
        launch_cost_order(ray_cost, w, n_rays, s);
    }
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        if (background) {
            // The reference's grid backward hands its final accum and log
            // transmittance to the background's (:1032-1064); here the grid's
            // own colours and transmittance come from a forward render.
            OptView ofg = o;
            ofg.bg = 0.f;
            hipLaunchKernelGGL((k_render_rays<DT, B, false>), dim3(blocks), dim3(kBlock), 0, s,
                               g, ofg, origins, dirs, ray_clip, n_rays, bg_buf,
                               bg_buf + 3 * n_rays, (int32_t*)nullptr, (uint32_t*)nullptr,
                               RecStream{}, (const uint32_t*)nullptr, ChunkLog{});
        }
        hipLaunchKernelGGL((k_render_backward<DT, B>), dim3(blocks), dim3(kBlock), 0, s, g, o,
                           origins, dirs, ray_clip, n_rays, grad_out, color_cache,
                           static_cast<const uint32_t*>(w.order), go);
    });
    if (st) return st;
    if (background &&
        (st = launch_background_backward(grid, opt, origins, dirs, n_rays, grad_out, color_cache,
                                         0, 1.f, bg_buf + 3 * n_rays, bg_buf, 1, 0.f, grads, s)))
        return st;
    return check_launch("volume_render_cuvol_backward");
}

extern "C" size_t svox2_ray_order_workspace_bytes(int64_t n_rays) {
    return n_rays > 0 ? order_workspace_bytes(n_rays) : 0;
}


// Whether the fused call records its samples for a batch of n rays (given a
// workspace of fused_workspace_bytes(n)).  Replaying records saves each ray's
// re-march; it pays up to about two ordering thresholds (measured on one
// MI355X: 0.780 against 0.855 ms at 65536 rays, 1.41 against 1.20 ms at
// 131072, where the records' traffic outweighs the re-march).  Larger
// batches re-march from the chunk log.  Slot indices must stay below 2^31
// (headers hold them as int32).
bool fused_records(int64_t n) {
    return (SVOX2_RECORD_ALL || n <= 2 * order_min_rays()) &&
           n * (kRecFixed + kRecOverflow) < INT32_MAX;
}

extern "C" size_t svox2_background_workspace_bytes(int64_t n_rays) {
    if (n_rays <= 0) return 0;
    return background_offset(n_rays) + 16 * static_cast<size_t>(n_rays);
}

extern "C" size_t svox2_volume_render_cuvol_fused_workspace_bytes(int64_t n_rays) {
    if (n_rays <= 0) return 0;
    return std::max(fused_records(n_rays) ? fused_workspace_bytes(n_rays)
                                          : logged_workspace_bytes(n_rays),
                    svox2_background_workspace_bytes(n_rays));
}

extern "C" int svox2_volume_render_cuvol_fused(const svox2_grid_t* grid,
                                               const float* origins, const float* dirs,
                                               const float* ray_clip, int64_t n_rays,
                                               const svox2_render_options_t* opt,
                                               const float* rgb_gt, float beta_loss,
                                               float sparsity_loss, float* rgb_out,
                                               float* log_transmit_out,
                                               const svox2_grads_t* grads, void* workspace,
                                               size_t workspace_bytes, void* stream) {
    int st = check_grid(grid, true);
    if (st) return st;
    OptView o;
    if ((st = make_opt(opt, o))) return st;
    if ((st = check_rays(origins, dirs, n_rays))) return st;
    GradOut go;
    if ((st = check_grads(grads, grid->sh_data_dim, go))) return st;
    if (n_rays == 0) return ok();
    SVOX2_REQUIRE(rgb_gt != nullptr && rgb_out != nullptr, "rgb_gt / rgb_out are NULL");
    SVOX2_REQUIRE(!(beta_loss > 0.f) || log_transmit_out != nullptr,
                  "beta_loss > 0 needs log_transmit_out (n floats)");
    if ((st = check_order_workspace(workspace, workspace_bytes, n_rays))) return st;
    const float norm = 2.f / static_cast<float>(3 * n_rays);
    const float beta = beta_loss / static_cast<float>(n_rays);
    const unsigned blocks = blocks_for(n_rays * kGroup, kBlock);
    float* lt = beta_loss > 0.f ? log_transmit_out : nullptr;
    // A background (render_lerp_kernel_cuvol.cu:1088-1150): the grid renders
    // with brightness 0, the background adds its colour (kept in the workspace
    // for its backward) before the grid's backward reads the total, and the
    // background's backward runs last.  Every ray re-marches: the records'
    // replay composes a ray's colour from the grid's samples alone.
    const bool background = grid_has_background(grid);
    float* bg_rgb = nullptr;
    if (background) {
        if ((st = check_background(grid, opt))) return st;
        SVOX2_REQUIRE(workspace != nullptr &&
                          workspace_bytes >= svox2_background_workspace_bytes(n_rays),
                      "a grid with a background needs a workspace of "
                      "svox2_background_workspace_bytes(n_rays) = %zu bytes",
                      svox2_background_workspace_bytes(n_rays));
        bg_rgb = background_buffer(workspace, n_rays);
        lt = log_transmit_out != nullptr ? log_transmit_out : bg_rgb + 3 * n_rays;
        o.bg = 0.f;
    }
    const hipStream_t s = as_stream(stream);
    OrderWorkspace w;
    if (workspace != nullptr && n_rays >= order_min_rays()) w = order_workspace(workspace, n_rays);
    RecStream rs{};
    ChunkLog log{};
    // Batches that re-march log the forward's chunk placements when the
    // workspace has room for them.
    if (workspace != nullptr && !background && !fused_records(n_rays) &&
        workspace_bytes >= logged_workspace_bytes(n_rays))
        log = chunk_log(workspace, n_rays);
    if (workspace != nullptr && !background && fused_records(n_rays) &&
        workspace_bytes >= fused_workspace_bytes(n_rays)) {
        rs = record_stream(workspace, n_rays);
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kSortBlock), 0, s,
                           reinterpret_cast<uint32_t*>(rs.next_block), 3 * kRecParts);
    }
    int bg_st = 0;
    st = dispatch(grid, [&](auto g, auto b, auto dtp) {
        constexpr int B = decltype(b)::value;
        using DT = std::remove_pointer_t<decltype(dtp)>;
        // Large batches: the forward runs in order of segment length, then
        // the backward in order of the forward's work counts (both sorts go
        // through w.cost / w.order; the forward rewrites w.cost by ray).
        if (w.order != nullptr) launch_length_order(grid, o, origins, dirs, ray_clip, n_rays, w, s);
        const uint32_t* order = w.order;
        if (rs.rec != nullptr) {
            hipLaunchKernelGGL((k_render_rays_rec<DT, B>), dim3(blocks), dim3(kBlock), 0, s, g,
                               o, origins, dirs, ray_clip, n_rays, rgb_out, lt, w.cost, rs,
                               order);
            hipLaunchKernelGGL((k_rec_colours<DT, B>),
                               dim3(blocks_for(std::min(rs.clist_cap, 2 * n_rays) * kGroup, kBlock)),
                               dim3(kBlock), 0, s, g, dirs, rs);
        }
        else if (log.rec != nullptr)
            hipLaunchKernelGGL((k_render_rays<DT, B, false, true>), dim3(blocks), dim3(kBlock), 0,
                               s, g, o, origins, dirs, ray_clip, n_rays, rgb_out, lt,
                               (int32_t*)nullptr, w.cost, rs, order, log);
        else
            hipLaunchKernelGGL((k_render_rays<DT, B, false>), dim3(blocks), dim3(kBlock), 0, s,
                               g, o, origins, dirs, ray_clip, n_rays, rgb_out, lt,
                               (int32_t*)nullptr, w.cost, rs, order, log);
        if (background)
            bg_st = launch_background(grid, opt, origins, dirs, nullptr, n_rays, lt, rgb_out,
                                      bg_rgb, s);
        // The key kernel cleared both histograms; this sort uses the second.
        if (w.order != nullptr) launch_cost_order(w.cost, w, n_rays, s, 1, true);
        if (rs.rec != nullptr)
            hipLaunchKernelGGL((k_render_backward_replay<DT, B>),
                               dim3(blocks_for((n_rays + rs.seg_cap) * kGroup,
                                               kBlock)),
                               dim3(kBlock), 0, s, g, dirs, n_rays, rgb_gt, rgb_out, lt, beta,
                               sparsity_loss, norm, static_cast<const uint32_t*>(w.order), go, rs);
        // Batches of up to 4 device fills use the 3-wave build.
        if (n_rays <= 4 * order_min_rays())
            hipLaunchKernelGGL((k_render_backward_mse<DT, B, SVOX2_BWD_WAVES_FEW>), dim3(blocks),
                               dim3(kBlock), 0, s, g, o, origins, dirs, ray_clip, n_rays, rgb_gt,
                               rgb_out, lt, beta, sparsity_loss, norm,
                               static_cast<const uint32_t*>(w.order), go, rs, log);
        else
            hipLaunchKernelGGL((k_render_backward_mse<DT, B, SVOX2_BWD_WAVES>), dim3(blocks),
                               dim3(kBlock), 0, s, g, o, origins, dirs, ray_clip, n_rays, rgb_gt,
                               rgb_out, lt, beta, sparsity_loss, norm,
                               static_cast<const uint32_t*>(w.order), go, rs, log);
    });
    if (st) return st;
    if (bg_st) return bg_st;
    if (background &&
        (st = launch_background_backward(grid, opt, origins, dirs, n_rays, rgb_gt, rgb_out, 1,
                                         norm, lt, bg_rgb, 0, sparsity_loss, grads, s)))
        return st;
    return check_launch("volume_render_cuvol_fused");
}

// ---------------------------------------------------------------------------
// The fused call without a caller workspace or any allocation: its workspace
// is a fixed device array in this library.  One copy of the array lives on
// every device the library runs on, at its own address, which
// hipGetSymbolAddress gives once per device; host code uses only that
// address.  Calls that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
namespace svox2amd {
namespace {

constexpr size_t kScratchBytes = size_t{40} << 20;  // 40 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up on its first use.
int scratch_ptr(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(SVOX2_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    SVOX2_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(SVOX2_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ok();
}

}  // namespace
}  // namespace svox2amd

extern "C" int svox2_volume_render_cuvol_fused_noalloc(const svox2_grid_t* grid,
                                                       const float* origins, const float* dirs,
                                                       const float* ray_clip, int64_t n_rays,
                                                       const svox2_render_options_t* opt,
                                                       const float* rgb_gt, float beta_loss,
                                                       float sparsity_loss, float* rgb_out,
                                                       float* log_transmit_out,
                                                       const svox2_grads_t* grads, void* stream) {
    SVOX2_REQUIRE(n_rays >= 0, "n_rays = %lld", static_cast<long long>(n_rays));
    const size_t need = svox2_volume_render_cuvol_fused_workspace_bytes(n_rays);
    SVOX2_REQUIRE(need <= kScratchBytes,
                  "%lld rays need %zu workspace bytes; the no-allocation fused call holds %zu "
                  "(8192 rays at least)",
                  static_cast<long long>(n_rays), need, kScratchBytes);
    void* ws = nullptr;
    if (n_rays > 0) {
        const int st = scratch_ptr(&ws);
        if (st) return st;
    }
    return svox2_volume_render_cuvol_fused(grid, origins, dirs, ray_clip, n_rays, opt, rgb_gt,
                                           beta_loss, sparsity_loss, rgb_out, log_transmit_out,
                                           grads, ws, n_rays > 0 ? kScratchBytes : 0, stream);
}
