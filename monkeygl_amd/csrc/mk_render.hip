// mk_render.hip -- VR / MIP / Surface ray casting on gfx950.
//
// Semantics follow the reference kernels d_render, d_renderMIP and
// d_renderSurface (core/kernel.cu:133-469): same ray setup, same adaptive
// step state machine with backtracking (getNextStep, kernel.cu:84-125), same
// compositing, shading and early exit. What changes is how the work maps to
// the hardware:
//   * explicit trilinear sampling of the bricked int16 volume in HBM
//     (no texture units);
//   * each ray jumps to its first possible in-VOI sample through the exact
//     float step lattice a_{k+1} = a_k + L1 (identical to the reference's
//     sequential accumulation) and stops after its last possible one;
//   * per-brick occupancy: a sample whose brick cannot produce nonzero
//     transfer-function alpha (VR), cannot raise the running maximum (MIP) or
//     cannot pass the 400 HU threshold (Surface) skips every fetch; the state
//     machine still runs with alpha 0, so results are unchanged;
//   * LUTs and the per-label window/alpha palette staged in LDS, per-label
//     opacity accumulators in LDS instead of a 65-entry private array;
//   * 16x16 pixel tiles of four 8x8 waves, mapped XCD-aware.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mk_context.h"

namespace {

constexpr int kMaxLdsBytes = 64 * 1024;
constexpr int kLutLdsEntries = 1024;  // stage LUTs in LDS up to 16 KiB

struct RenderArgs {
    MkVolView v;
    const float4* lut;
    MkLutTable tab;
    const uint8_t* occ;    // per 8^3 brick (VR, Surface)
    const float* mipb;     // per 8^3 brick (MIP)
    const uint8_t* occ2;   // per 32^3 macro cell (4x4x4 bricks)
    const uint8_t* blabel; // per brick: its only label, or 0xFF when mixed
    const float* mipb2;
    int nmx, nmy;          // macro cell dims (x, y)
    const float* steps;
    int nsteps;
    int nlut_lds;      // entries staged in LDS (0 = read LUTs from global)
    int nslot;         // per-label accumulator slots (masked VR)
    uint8_t slot_of[MK_NUM_LABELS + 3];
    uint8_t* out;
    uint8_t* touch_v;  // footprint maps (measurement variant only)
    uint8_t* touch_m;
    int width, height;
    int invert_z;
    float mls[3];
    float sv[3];
    float xT, yT;
    double scale;
    float M[9];
    mk_voi voi;
    float bg[4];
    float l1;
    float exit_alpha;
    int max_iter;      // hard bound on loop trips per ray
};

struct AwTable {
    float alpha[MK_NUM_LABELS];
    float ww[MK_NUM_LABELS];
    float wl[MK_NUM_LABELS];
};

__device__ inline void ray_pos(const RenderArgs& A, float u, float v, float acc, float p[3]) {
    float fy = (float)(((double)acc - 0.866) * A.scale);
    float x = u * A.M[0] + fy * A.M[1] + v * A.M[2];
    float y = u * A.M[3] + fy * A.M[4] + v * A.M[5];
    float z = u * A.M[6] + fy * A.M[7] + v * A.M[8];
    p[0] = x * A.mls[0] + 0.5f;
    p[1] = y * A.mls[1] + 0.5f;
    p[2] = z * A.mls[2] + 0.5f;
    if (A.invert_z) p[2] = 1.0f - p[2];
}

// Linear model of a ray, pos(a) ~ p0 + a * d, in normalized volume units.
struct RayGeom {
    float p0[3];
    float d[3];
};

// Conservative [a_in, a_out] of accumulated lengths whose samples can pass
// the VOI test; returns false when none can.
__device__ inline bool ray_interval(const RenderArgs& A, float u, float v, RayGeom& g, float& a_in, float& a_out) {
    ray_pos(A, u, v, 0.0f, g.p0);
    // pos(a) at a = 0 is taken at fy = -0.866*scale; shift to the a origin
    g.d[0] = A.M[1] * (float)A.scale * A.mls[0];
    g.d[1] = A.M[4] * (float)A.scale * A.mls[1];
    g.d[2] = A.M[7] * (float)A.scale * A.mls[2];
    if (A.invert_z) g.d[2] = -g.d[2];
    const int lo_i[3] = {A.voi.left, A.voi.posterior, A.voi.head};
    const int hi_i[3] = {A.voi.right, A.voi.anterior, A.voi.foot};
    const int n[3] = {A.v.nx, A.v.ny, A.v.nz};
    float t0 = 0.0f, t1 = 1.7321f;
    const float eps = 1e-3f;
    for (int i = 0; i < 3; ++i) {
        // trunc(p*N) in [lo, hi]  =>  p in (lo-1, hi+1)/N
        float lo = ((float)lo_i[i] - 1.0f) / (float)n[i] - eps;
        float hi = ((float)hi_i[i] + 1.0f) / (float)n[i] + eps;
        if (fabsf(g.d[i]) < 1e-12f) {
            if (!(g.p0[i] >= lo && g.p0[i] <= hi)) return false;
        } else {
            float ta = (lo - g.p0[i]) / g.d[i], tb = (hi - g.p0[i]) / g.d[i];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
        }
    }
    const float m = 4.0f * A.l1;
    a_in = fmaxf(t0 - m, 0.0f);
    a_out = t1 + m;
    return a_in <= a_out;
}

// Accumulated length below which the ray is certain to stay in the cell of
// 2^shift voxels per axis that holds voxel (cx, cy, cz): the earliest exit
// over the three axes, with the cell shrunk by a margin that covers the
// rounding of the sample positions.
__device__ inline float cell_exit(const RenderArgs& A, const RayGeom& g, int cx, int cy, int cz, int shift) {
    const int c[3] = {cx >> shift, cy >> shift, cz >> shift};
    const float n[3] = {(float)A.v.nx, (float)A.v.ny, (float)A.v.nz};
    const float size = (float)(1 << shift);
    const float margin = 1e-5f;
    float t = INFINITY;
    for (int i = 0; i < 3; ++i) {
        if (g.d[i] > 1e-12f) {
            float face = ((float)(c[i] + 1) * size) / n[i] - margin;
            t = fminf(t, (face - g.p0[i]) / g.d[i]);
        } else if (g.d[i] < -1e-12f) {
            float face = ((float)c[i] * size) / n[i] + margin;
            t = fminf(t, (face - g.p0[i]) / g.d[i]);
        }
    }
    return t;
}

__device__ inline bool in_voi(const RenderArgs& A, const float p[3], int& ix, int& iy, int& iz) {
    float fx = p[0] * (float)A.v.nx, fy = p[1] * (float)A.v.ny, fz = p[2] * (float)A.v.nz;
    fx = fminf(fmaxf(fx, -1e9f), 1e9f);
    fy = fminf(fmaxf(fy, -1e9f), 1e9f);
    fz = fminf(fmaxf(fz, -1e9f), 1e9f);
    ix = (int)fx; iy = (int)fy; iz = (int)fz;
    return !(ix < A.voi.left || ix > A.voi.right || iy < A.voi.posterior || iy > A.voi.anterior || iz < A.voi.head ||
             iz > A.voi.foot);
}

__device__ inline uint32_t macro_index(const RenderArgs& A, int cx, int cy, int cz) {
    return ((uint32_t)(cz >> 5) * (uint32_t)A.nmy + (uint32_t)(cy >> 5)) * (uint32_t)A.nmx + (uint32_t)(cx >> 5);
}

// tex1D linear/clamp of label's LUT at normalized s.
template <bool LDS>
__device__ inline float4 lut_sample(const RenderArgs& A, const float4* s_lut, int label, float s) {
    int n = A.tab.len[label];
    if (n <= 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    int off = A.tab.off[label];
    s = s >= -1.0f ? s : -1.0f;
    int i0, i1;
    float a;
    mk_split(s, n, i0, i1, a);
    float4 c0 = LDS ? s_lut[off + i0] : A.lut[off + i0];
    float4 c1 = LDS ? s_lut[off + i1] : A.lut[off + i1];
    return make_float4(c0.x + a * (c1.x - c0.x), c0.y + a * (c1.y - c0.y), c0.z + a * (c1.z - c0.z),
                       c0.w + a * (c1.w - c0.w));
}

__device__ inline float3 light_dir(const RenderArgs& A) {
    float3 l = make_float3(A.M[1], A.M[4], A.M[7]);
    float r = rsqrtf(l.x * l.x + l.y * l.y + l.z * l.z);
    return make_float3(l.x * r, l.y * r, l.z * r);
}

// Normalized central-difference gradient one voxel apart (kernel.cu:60-67);
// a zero gradient gives NaN, which fails every "> 0" test downstream. One
// axis per trip keeps 16 loads in flight instead of 48.
__device__ inline float3 gradient(const RenderArgs& A, const float p[3]) {
    float N[3];
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        float q0[3] = {p[0], p[1], p[2]}, q1[3] = {p[0], p[1], p[2]};
        q0[i] += A.sv[i];
        q1[i] -= A.sv[i];
        N[i] = mk_sample_hu(A.v, q0[0], q0[1], q0[2]) - mk_sample_hu(A.v, q1[0], q1[1], q1[2]);
    }
    if (A.invert_z) N[2] = -N[2];
    float r = rsqrtf(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    return make_float3(N[0] * r, N[1] * r, N[2] * r);
}

// Footprint marking for the measurement variant: every voxel the reference
// rule reads is set to 1 in a dense x-fastest byte map.
__device__ inline void touch_footprint(const RenderArgs& A, float fx, float fy, float fz) {
    int x0, x1, y0, y1, z0, z1;
    float a;
    mk_split(fx, A.v.nx, x0, x1, a);
    mk_split(fy, A.v.ny, y0, y1, a);
    mk_split(fz, A.v.nz, z0, z1, a);
    const size_t nx = A.v.nx, nxy = (size_t)A.v.nx * A.v.ny;
    const int xs[2] = {x0, x1}, ys[2] = {y0, y1}, zs[2] = {z0, z1};
    for (int k = 0; k < 8; ++k) A.touch_v[zs[k >> 2] * nxy + ys[(k >> 1) & 1] * nx + xs[k & 1]] = 1;
}

__device__ inline void touch_gradient(const RenderArgs& A, const float p[3]) {
    touch_footprint(A, p[0] + A.sv[0], p[1], p[2]);
    touch_footprint(A, p[0] - A.sv[0], p[1], p[2]);
    touch_footprint(A, p[0], p[1] + A.sv[1], p[2]);
    touch_footprint(A, p[0], p[1] - A.sv[1], p[2]);
    touch_footprint(A, p[0], p[1], p[2] + A.sv[2]);
    touch_footprint(A, p[0], p[1], p[2] - A.sv[2]);
}

__device__ inline void touch_mask(const RenderArgs& A, int cx, int cy, int cz) {
    if (A.touch_m) A.touch_m[((size_t)cz * A.v.ny + cy) * A.v.nx + cx] = 1;
}

__device__ inline void store_rgb(const RenderArgs& A, int x, int y, float4 c) {
    uint8_t* o = A.out + 3 * ((size_t)y * A.width + x);
    o[0] = (uint8_t)(unsigned)(mk_sat(c.x) * 255.0f);
    o[1] = (uint8_t)(unsigned)(mk_sat(c.y) * 255.0f);
    o[2] = (uint8_t)(unsigned)(mk_sat(c.z) * 255.0f);
}

__device__ inline float4 bg4(const RenderArgs& A) { return make_float4(A.bg[0], A.bg[1], A.bg[2], A.bg[3]); }

template <int TW>
__device__ inline bool tile_pixel(const RenderArgs& A, int& x, int& y) {
    const uint32_t tiles_x = (A.width + TW - 1) / TW;
    const uint32_t tiles = tiles_x * ((A.height + TW - 1) / TW);
    uint32_t t = mk_xcd_tile(blockIdx.x, tiles);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int WPR = TW / 8;  // waves per tile row
    x = (int)(t % tiles_x) * TW + (wave % WPR) * 8 + (lane & 7);
    y = (int)(t / tiles_x) * TW + (wave / WPR) * 8 + (lane >> 3);
    return x < A.width && y < A.height;
}

// Shared-memory carve-up common to the kernels.
struct Smem {
    float4* aw;     // MK_NUM_LABELS x (alpha, ww, wl, -)
    float4* lut;    // nlut_lds entries
    float* aobj;    // nslot x blockDim
};

__device__ inline Smem stage_shared(const RenderArgs& A, const AwTable& T, bool stage_lut) {
    extern __shared__ float4 smem[];
    Smem s;
    s.aw = smem;
    s.lut = smem + 68;
    s.aobj = (float*)(s.lut + A.nlut_lds);
    for (int i = threadIdx.x; i < MK_NUM_LABELS; i += blockDim.x) s.aw[i] = make_float4(T.alpha[i], T.ww[i], T.wl[i], 0.f);
    if (stage_lut)
        for (int i = threadIdx.x; i < A.nlut_lds; i += blockDim.x) s.lut[i] = A.lut[i];
    __syncthreads();
    return s;
}

__device__ inline float start_acc(const RenderArgs& A, float a_in) {
    int k0 = (int)(a_in / A.l1) - 2;
    k0 = k0 < 0 ? 0 : (k0 >= A.nsteps ? A.nsteps - 1 : k0);
    return A.steps[k0];
}

// Advances acc over the lattice points a_{k+1} = a_k + L1 that lie below
// a_skip (all inside one cell that cannot change the result); it counts trips
// against the same budget as the marching loop.
__device__ inline void skip_lattice(float& acc, float l1, float a_skip, int& it, int max_iter) {
    while (acc + l1 < a_skip && it < max_iter) {
        acc += l1;
        ++it;
    }
}

// d_render (core/kernel.cu:133-264). TOUCH: measurement variant that skips
// nothing and marks every voxel the reference rule reads.
template <int TW, bool MASK, bool LDS_LUT, bool TOUCH>
__global__ void __launch_bounds__(TW * TW) k_vr(RenderArgs A, AwTable T) {
    Smem s = stage_shared(A, T, LDS_LUT);
    int x, y;
    if (!tile_pixel<TW>(A, x, y)) return;
    const int tid = threadIdx.x;
    constexpr int BS = TW * TW;

    const float u = ((float)x - (float)A.width / 2.0f - A.xT) / (float)A.height;
    const float v = ((float)y - (float)A.height / 2.0f - A.yT) / (float)A.height;
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
    float a_in, a_out;
    RayGeom g;
    if (ray_interval(A, u, v, g, a_in, a_out)) {
        const float L1 = A.l1, L4 = L1 / 4.0f, L8 = L1 / 8.0f;
        float step = L1, aPre = 0.f, aAcc = 0.f, aobj0 = 0.f;
        if (MASK)
            for (int k = 0; k < A.nslot; ++k) s.aobj[k * BS + tid] = 0.f;
        float acc = start_acc(A, a_in);
        // skip-table entries of the brick the ray is in, reloaded when it
        // moves to another brick
        uint32_t cur_brick = 0xFFFFFFFFu;
        int cur_macro_empty = 0, cur_occ = 1, cur_label = 0xFF;
        for (int it = 0; it < A.max_iter && (double)acc < 1.732 && acc <= a_out; ++it) {
            float p[3];
            ray_pos(A, u, v, acc, p);
            int ix, iy, iz;
            if (!in_voi(A, p, ix, iy, iz)) { acc += step; continue; }
            const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1), cz = mk_clampi(iz, 0, A.v.nz - 1);
            // a sample in an empty cell reads alpha 0 exactly (the skip tables
            // are conservative), so it needs no fetch
            int empty_shift = 0;
            if (!TOUCH) {
                const uint32_t b = mk_brick_index(A.v, cx, cy, cz);
                if (b != cur_brick) {
                    cur_brick = b;
                    cur_macro_empty = !A.occ2[macro_index(A, cx, cy, cz)];
                    cur_occ = A.occ[b];
                    cur_label = MASK ? A.blabel[b] : 0;
                }
                empty_shift = cur_macro_empty ? 5 : (cur_occ ? 0 : 3);
            }
            float a = 0.f;
            float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
            int label = 0;
            if (!empty_shift) {
                if (MASK) label = (TOUCH || cur_label == 0xFF) ? A.v.mask[mk_voxel_offset(A.v, cx, cy, cz)] : cur_label;
                if (TOUCH) {
                    if (MASK) touch_mask(A, cx, cy, cz);
                    touch_footprint(A, p[0], p[1], p[2]);
                }
                const float4 aw = s.aw[label];
                float t = mk_sample_hu(A.v, p[0], p[1], p[2]);
                t = (t - aw.z) / aw.y + 0.5f;
                t = t > 1.0f ? 1.0f : t;
                col = lut_sample<LDS_LUT>(A, s.lut, label, t);
                a = col.w;
            }
            // getNextStep (kernel.cu:84-125)
            if (a > 0.f && step != L1) a = 1.0f - __builtin_amdgcn_exp2f((step == L4 ? 0.25f : 0.125f) * __builtin_amdgcn_logf(1.0f - a));
            if (acc > 0.f) {
                if (fmaxf(a, aPre) > 0.001f) {
                    if (step == L1) { acc -= (L1 - L4); step = L4; continue; }
                    if (step == L4) { acc -= (L4 - L8); step = L8; continue; }
                } else {
                    step = (step == L8) ? L4 : L1;
                }
            }
            aPre = a;
            acc += step;
            if (a > 0.0005f) {
                const float cap = s.aw[label].x;
                float* slot = MASK ? &s.aobj[A.slot_of[label] * BS + tid] : &aobj0;
                if (*slot < cap) {
                    // tracing (kernel.cu:49-81)
                    if (TOUCH) touch_gradient(A, p);
                    const float3 N = gradient(A, p);
                    const float3 L = light_dir(A);
                    const float d = N.x * L.x + N.y * L.y + N.z * L.z;
                    float k = 0.35f;
                    if (d > 0.f) {
                        float d2 = d * d, d4 = d2 * d2;
                        k += d * 0.8f + 0.16f * (d4 * d4);
                    }
                    const float w = (1.0f - aAcc) * a;
                    sum.x += w * (col.x * k);
                    sum.y += w * (col.y * k);
                    sum.z += w * (col.z * k);
                    sum.w += w * (a * k);
                    *slot += w;
                    aAcc += w;
                }
            }
            if (aAcc > A.exit_alpha) break;
            // steady state in an empty cell: every further lattice point inside
            // it is accepted with alpha 0 and changes nothing but acc
            if (empty_shift && step == L1) skip_lattice(acc, L1, cell_exit(A, g, cx, cy, cz, empty_shift), it, A.max_iter);
        }
    }
    if (sum.x == 0.f && sum.y == 0.f && sum.z == 0.f && sum.w == 0.f) sum = bg4(A);
    store_rgb(A, x, y, sum);
}

// d_renderMIP (core/kernel.cu:266-364).
template <int TW, bool MASK, bool TOUCH>
__global__ void __launch_bounds__(TW * TW) k_mip(RenderArgs A, AwTable T) {
    Smem s = stage_shared(A, T, false);
    int x, y;
    if (!tile_pixel<TW>(A, x, y)) return;
    const float u = ((float)x - (float)A.width / 2.0f - A.xT) / (float)A.height;
    const float v = ((float)y - (float)A.height / 2.0f - A.yT) / (float)A.height;
    float aAcc = 0.f;
    float a_in, a_out;
    RayGeom g;
    if (ray_interval(A, u, v, g, a_in, a_out)) {
        const float L1 = A.l1;
        float acc = start_acc(A, a_in);
        uint32_t cur_brick = 0xFFFFFFFFu;
        float cur_bm = INFINITY, cur_bb = INFINITY;
        int cur_label = 0xFF;
        for (int it = 0; it < A.max_iter && (double)acc < 1.732 && acc <= a_out; ++it) {
            float p[3];
            ray_pos(A, u, v, acc, p);
            int ix, iy, iz;
            if (in_voi(A, p, ix, iy, iz)) {
                const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1),
                          cz = mk_clampi(iz, 0, A.v.nz - 1);
                // a cell whose bound cannot exceed the running maximum leaves it unchanged
                int skip_shift = 0;
                if (!TOUCH) {
                    const uint32_t b = mk_brick_index(A.v, cx, cy, cz);
                    if (b != cur_brick) {
                        cur_brick = b;
                        cur_bm = A.mipb2[macro_index(A, cx, cy, cz)];
                        cur_bb = A.mipb[b];
                        cur_label = MASK ? A.blabel[b] : 0;
                    }
                    skip_shift = cur_bm <= aAcc ? 5 : (cur_bb <= aAcc ? 3 : 0);
                }
                if (!skip_shift) {
                    int label = MASK ? ((TOUCH || cur_label == 0xFF) ? A.v.mask[mk_voxel_offset(A.v, cx, cy, cz)] : cur_label) : 0;
                    if (TOUCH && MASK) touch_mask(A, cx, cy, cz);
                    const float4 aw = s.aw[label];
                    if (aw.x > 0.f) {
                        if (TOUCH) touch_footprint(A, p[0], p[1], p[2]);
                        float t = mk_sample_hu(A.v, p[0], p[1], p[2]);
                        t = (t - aw.z) / aw.y + 0.5f;
                        if (aAcc < t) aAcc = t;
                    }
                }
                acc += L1;
                if (skip_shift) skip_lattice(acc, L1, cell_exit(A, g, cx, cy, cz, skip_shift), it, A.max_iter);
            } else {
                acc += L1;
            }
        }
    }
    float4 c = aAcc <= 0.f ? bg4(A) : make_float4(aAcc, aAcc, aAcc, aAcc);
    store_rgb(A, x, y, c);
}

// d_renderSurface (core/kernel.cu:366-469).
template <int TW, bool TOUCH>
__global__ void __launch_bounds__(TW * TW) k_surface(RenderArgs A, AwTable T) {
    (void)stage_shared(A, T, false);
    int x, y;
    if (!tile_pixel<TW>(A, x, y)) return;
    const float u = ((float)x - (float)A.width / 2.0f - A.xT) / (float)A.height;
    const float v = ((float)y - (float)A.height / 2.0f - A.yT) / (float)A.height;
    float aAcc = 0.f;
    float a_in, a_out;
    RayGeom g;
    if (ray_interval(A, u, v, g, a_in, a_out)) {
        const float L1 = A.l1;
        float acc = start_acc(A, a_in);
        uint32_t cur_brick = 0xFFFFFFFFu;
        int cur_macro_empty = 0, cur_occ = 1;
        for (int it = 0; it < A.max_iter && (double)acc < 1.732 && acc <= a_out; ++it) {
            float p[3];
            ray_pos(A, u, v, acc, p);
            int ix, iy, iz;
            if (in_voi(A, p, ix, iy, iz)) {
                const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1),
                          cz = mk_clampi(iz, 0, A.v.nz - 1);
                int skip_shift = 0;
                if (!TOUCH) {
                    const uint32_t b = mk_brick_index(A.v, cx, cy, cz);
                    if (b != cur_brick) {
                        cur_brick = b;
                        cur_macro_empty = !A.occ2[macro_index(A, cx, cy, cz)];
                        cur_occ = A.occ[b];
                    }
                    skip_shift = cur_macro_empty ? 5 : (cur_occ ? 0 : 3);
                }
                if (!skip_shift) {
                    if (TOUCH) touch_footprint(A, p[0], p[1], p[2]);
                    float t = mk_sample_hu(A.v, p[0], p[1], p[2]);
                    if (t > 400.0f) {
                        if (TOUCH) touch_gradient(A, p);
                        const float3 N = gradient(A, p);
                        const float3 L = light_dir(A);
                        aAcc = N.x * L.x + N.y * L.y + N.z * L.z;
                        if (aAcc > 0.f) break;
                    }
                }
                acc += L1;
                if (skip_shift) skip_lattice(acc, L1, cell_exit(A, g, cx, cy, cz, skip_shift), it, A.max_iter);
            } else {
                acc += L1;
            }
        }
    }
    float4 c;
    if (aAcc <= 0.f) c = bg4(A);
    else {
        float g2 = (float)((double)aAcc + 0.05);
        c = make_float4(g2, g2, g2, 1.0f);
    }
    store_rgb(A, x, y, c);
}

// Per-brick skip table for one frame. VR: 1 when some label present in the
// brick can read a LUT entry with nonzero alpha over the brick's value range
// (with margin). MIP: upper bound of the windowed value over labels with
// alpha > 0. Surface: 1 when the brick's maximum can exceed 400 HU.
__global__ void __launch_bounds__(256) k_occupancy(const int16_t* __restrict__ vmin, const int16_t* __restrict__ vmax,
                                                    const uint64_t* __restrict__ lab_lo, const uint8_t* __restrict__ lab_hi,
                                                    uint32_t nb, int type, AwTable T, MkLutTable tab,
                                                    const int* __restrict__ lutnz, uint8_t* __restrict__ occ,
                                                    float* __restrict__ mipb) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const float k = 32768.0f / 32767.0f;
    float lo = (float)vmin[b] * k, hi = (float)vmax[b] * k;
    lo -= 1.0f + fabsf(lo) * 1e-5f;
    hi += 1.0f + fabsf(hi) * 1e-5f;
    if (type == MK_RENDER_SURFACE) {
        occ[b] = hi > 400.0f;
        return;
    }
    uint64_t bits = lab_lo[b];
    bool has64 = lab_hi[b] != 0;
    float bound = -INFINITY;
    uint8_t any = 0;
    for (int pass = 0; pass < 2; ++pass) {
        uint64_t w = pass == 0 ? bits : (has64 ? 1ull : 0ull);
        while (w) {
            int l = __ffsll((long long)w) - 1;
            w &= w - 1;
            if (pass == 1) l = 64;
            const float ww = T.ww[l], wl = T.wl[l];
            float s0 = (lo - wl) / ww, s1 = (hi - wl) / ww;
            float smin = fminf(s0, s1) + 0.5f, smax = fmaxf(s0, s1) + 0.5f;
            bool finite = isfinite(smin) && isfinite(smax);
            if (type == MK_RENDER_MIP) {
                if (T.alpha[l] > 0.f) {
                    float bnd = finite ? smax + 1e-4f + fabsf(smax) * 1e-5f : INFINITY;
                    bound = fmaxf(bound, bnd);
                }
                continue;
            }
            const int n = tab.len[l];
            if (n <= 0) continue;
            int i0 = 0, i1 = n - 1;
            if (finite) {
                smin = fminf(smin, 1.0f);
                smax = fminf(smax, 1.0f);
                float f0 = floorf(smin * (float)n - 0.5f) - 1.0f;
                float f1 = floorf(smax * (float)n - 0.5f) + 2.0f;
                i0 = (int)fminf(fmaxf(f0, 0.0f), (float)(n - 1));
                i1 = (int)fminf(fmaxf(f1, 0.0f), (float)(n - 1));
            }
            const int* p = lutnz + tab.off[l] + l;
            if (p[i1 + 1] - p[i0] > 0) any = 1;
        }
    }
    if (type == MK_RENDER_MIP) mipb[b] = bound;
    else occ[b] = any;
}

// Macro cells of 4x4x4 bricks: OR of the brick occupancy, max of the MIP bound.
__global__ void __launch_bounds__(256) k_macro(const uint8_t* __restrict__ occ, const float* __restrict__ mipb,
                                                int nbx, int nby, int nbz, int nmx, int nmy, int nmz, int type,
                                                uint8_t* __restrict__ occ2, float* __restrict__ mipb2) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (uint32_t)(nmx * nmy * nmz)) return;
    const int mx = m % nmx, my = (m / nmx) % nmy, mz = m / (nmx * nmy);
    uint8_t any = 0;
    float bound = -INFINITY;
    for (int dz = 0; dz < 4; ++dz)
        for (int dy = 0; dy < 4; ++dy)
            for (int dx = 0; dx < 4; ++dx) {
                const int bx = mx * 4 + dx, by = my * 4 + dy, bz = mz * 4 + dz;
                if (bx >= nbx || by >= nby || bz >= nbz) continue;
                const uint32_t b = ((uint32_t)bz * nby + by) * nbx + bx;
                if (type == MK_RENDER_MIP) bound = fmaxf(bound, mipb[b]);
                else any |= occ[b];
            }
    if (type == MK_RENDER_MIP) mipb2[m] = bound;
    else occ2[m] = any;
}

mk_status ensure_occupancy(mk_context* ctx, const mk_render_params* p, const AwTable& T) {
    bool same = ctx->occ_valid && ctx->occ_type == p->render_type && ctx->occ_lut_version == ctx->lut_version &&
                ctx->occ_vol_version == ctx->vol_version &&
                memcmp(ctx->occ_aw, p->alpha_wwwl, sizeof(ctx->occ_aw)) == 0;
    if (same) return MK_OK;
    uint32_t nb = (uint32_t)ctx->nbricks();
    hipLaunchKernelGGL(k_occupancy, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, ctx->stats.vmin, ctx->stats.vmax,
                       ctx->stats.lab_lo, ctx->stats.lab_hi, nb, p->render_type, T, ctx->lut_tab, ctx->d_lutnz, ctx->d_occ,
                       ctx->d_mipb);
    MK_HIP(ctx, hipGetLastError());
    const int nmx = (ctx->nbx + 3) / 4, nmy = (ctx->nby + 3) / 4, nmz = (ctx->nbz + 3) / 4;
    const uint32_t nm = (uint32_t)(nmx * nmy * nmz);
    hipLaunchKernelGGL(k_macro, dim3((nm + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_occ, ctx->d_mipb, ctx->nbx,
                       ctx->nby, ctx->nbz, nmx, nmy, nmz, p->render_type, ctx->d_occ2, ctx->d_mipb2);
    MK_HIP(ctx, hipGetLastError());
    ctx->occ_valid = true;
    ctx->occ_type = p->render_type;
    ctx->occ_lut_version = ctx->lut_version;
    ctx->occ_vol_version = ctx->vol_version;
    memcpy(ctx->occ_aw, p->alpha_wwwl, sizeof(ctx->occ_aw));
    return MK_OK;
}

mk_status validate(mk_context* ctx, const mk_render_params* p) {
    if (!p) return mk_fail(ctx, MK_ERR_ARG, "render params are null");
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "render before any volume was uploaded");
    if (p->width <= 0 || p->height <= 0 || p->width > 16384 || p->height > 16384)
        return mk_fail(ctx, MK_ERR_ARG, "bad image size %dx%d", p->width, p->height);
    if (p->render_type < MK_RENDER_VR || p->render_type > MK_RENDER_SURFACE)
        return mk_fail(ctx, MK_ERR_ARG, "unknown render type %d", p->render_type);
    if (!(p->step_scale > 0.0f) || !std::isfinite(p->step_scale) || ctx->nz / p->step_scale > 1e5f)
        return mk_fail(ctx, MK_ERR_ARG, "bad step_scale %g", (double)p->step_scale);
    return MK_OK;
}

mk_status launch_render(mk_context* ctx, const mk_render_params* p, uint8_t* dev_rgb, void* stream,
                        uint8_t* touch_v, uint8_t* touch_m) {
    mk_status st = validate(ctx, p);
    if (st) return st;
    if (!dev_rgb) return mk_fail(ctx, MK_ERR_ARG, "null output");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    st = mk_upload_luts(ctx);
    if (st) return st;
    AwTable T;
    for (int l = 0; l < MK_NUM_LABELS; ++l) {
        T.alpha[l] = p->alpha_wwwl[l].alpha;
        T.ww[l] = p->alpha_wwwl[l].ww;
        T.wl[l] = p->alpha_wwwl[l].wl;
    }
    st = ensure_occupancy(ctx, p, T);
    if (st) return st;
    const float l1 = 1.0f / (float)ctx->nz * p->step_scale;
    st = mk_ensure_steps(ctx, l1);
    if (st) return st;

    RenderArgs A;
    memset(&A, 0, sizeof(A));
    A.v = ctx->view();
    A.lut = ctx->d_lut;
    A.tab = ctx->lut_tab;
    A.occ = ctx->d_occ;
    A.mipb = ctx->d_mipb;
    A.occ2 = ctx->d_occ2;
    A.blabel = ctx->stats.blabel;
    A.mipb2 = ctx->d_mipb2;
    A.nmx = (ctx->nbx + 3) / 4;
    A.nmy = (ctx->nby + 3) / 4;
    A.steps = ctx->d_steps;
    A.nsteps = ctx->steps_n;
    A.out = dev_rgb;
    A.touch_v = touch_v;
    A.touch_m = touch_m;
    A.width = p->width;
    A.height = p->height;
    A.invert_z = p->invert_z != 0;
    for (int i = 0; i < 3; ++i) {
        A.mls[i] = p->max_len_spacing[i];
        A.sv[i] = p->spacing_voxel[i];
    }
    A.xT = p->x_translate;
    A.yT = p->y_translate;
    A.scale = (double)p->scale;
    memcpy(A.M, p->transform, sizeof(A.M));
    A.voi = p->voi;
    memcpy(A.bg, p->background, sizeof(A.bg));
    A.l1 = l1;
    A.exit_alpha = p->exit_alpha;
    // each accepted sample advances by at least L1/8 and at most two rejections
    // precede an acceptance: 3 * 8 trips per base step bound every ray
    A.max_iter = 24 * ctx->steps_n + 64;

    const bool mask = ctx->d_mask != nullptr && p->render_type != MK_RENDER_SURFACE;
    int nslot = 0;
    for (int l = 0; l < MK_NUM_LABELS; ++l) {
        bool used = l < 64 ? ((ctx->used_lo >> l) & 1ull) != 0 : ctx->used_hi;
        A.slot_of[l] = used ? (uint8_t)nslot : 0;
        if (used) ++nslot;
    }
    A.nslot = mask ? nslot : 0;
    int total_lut = 0;
    for (int l = 0; l < MK_NUM_LABELS; ++l)
        total_lut = std::max(total_lut, ctx->lut_tab.off[l] + ctx->lut_tab.len[l]);
    const bool lds_lut = p->render_type == MK_RENDER_VR && total_lut <= kLutLdsEntries;
    A.nlut_lds = lds_lut ? total_lut : 0;

    int tw = 16;
    size_t lds = 68 * sizeof(float4) + (size_t)A.nlut_lds * sizeof(float4) + (size_t)A.nslot * 256 * sizeof(float);
    if (lds > kMaxLdsBytes) {
        tw = 8;
        lds = 68 * sizeof(float4) + (size_t)A.nlut_lds * sizeof(float4) + (size_t)A.nslot * 64 * sizeof(float);
        if (lds > kMaxLdsBytes) return mk_fail(ctx, MK_ERR_ARG, "too many labels for the LDS budget");
    }
    const uint32_t tiles = (uint32_t)(((p->width + tw - 1) / tw) * ((p->height + tw - 1) / tw));
    hipStream_t s = (hipStream_t)stream;
    // the skip tables were (re)built on the context stream
    st = mk_order_after_context(ctx, s);
    if (st) return st;
    dim3 grid(tiles), block(tw * tw);
#define MK_LAUNCH(K) hipLaunchKernelGGL(K, grid, block, lds, s, A, T)
#define MK_VR(TOUCH)                                                                              \
    if (tw == 16) {                                                                               \
        if (mask) { if (lds_lut) MK_LAUNCH((k_vr<16, true, true, TOUCH>)); else MK_LAUNCH((k_vr<16, true, false, TOUCH>)); } \
        else { if (lds_lut) MK_LAUNCH((k_vr<16, false, true, TOUCH>)); else MK_LAUNCH((k_vr<16, false, false, TOUCH>)); } \
    } else {                                                                                      \
        if (lds_lut) MK_LAUNCH((k_vr<8, true, true, TOUCH>)); else MK_LAUNCH((k_vr<8, true, false, TOUCH>)); \
    }
    const bool touch = touch_v != nullptr;
    if (p->render_type == MK_RENDER_VR) {
        if (touch) { MK_VR(true) } else { MK_VR(false) }
    } else if (p->render_type == MK_RENDER_MIP) {
        if (touch) { if (mask) MK_LAUNCH((k_mip<16, true, true>)); else MK_LAUNCH((k_mip<16, false, true>)); }
        else { if (mask) MK_LAUNCH((k_mip<16, true, false>)); else MK_LAUNCH((k_mip<16, false, false>)); }
    } else {
        if (touch) MK_LAUNCH((k_surface<16, true>)); else MK_LAUNCH((k_surface<16, false>));
    }
#undef MK_VR
#undef MK_LAUNCH
    MK_HIP(ctx, hipGetLastError());
    return MK_OK;
}

}  // namespace

mk_status mk_ensure_steps(mk_context* ctx, float l1) {
    if (ctx->steps_l1 == l1 && ctx->d_steps) return MK_OK;
    std::vector<float> a;
    a.reserve((size_t)(1.75 / l1) + 8);
    float acc = 0.0f;
    a.push_back(acc);
    while ((double)acc < 1.74) {
        acc += l1;
        a.push_back(acc);
    }
    mk_status st = mk_ensure_buffer(ctx, (void**)&ctx->d_steps, &ctx->steps_cap, a.size() * sizeof(float));
    if (st) return st;
    MK_HIP(ctx, hipMemcpyAsync(ctx->d_steps, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->steps_l1 = l1;
    ctx->steps_n = (int)a.size();
    return MK_OK;
}

extern "C" mk_status mk_render_async(mk_context* ctx, const mk_render_params* p, uint8_t* dev_rgb, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    return launch_render(ctx, p, dev_rgb, stream, nullptr, nullptr);
}

extern "C" mk_status mk_render_footprint(mk_context* ctx, const mk_render_params* p, uint8_t* dev_rgb,
                                         uint8_t* dev_touch_vol, uint8_t* dev_touch_mask, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev_touch_vol) return mk_fail(ctx, MK_ERR_ARG, "null footprint map");
    return launch_render(ctx, p, dev_rgb, stream, dev_touch_vol, ctx->d_mask ? dev_touch_mask : nullptr);
}

extern "C" mk_status mk_render(mk_context* ctx, const mk_render_params* p, uint8_t* host_rgb) {
    if (!ctx) return MK_ERR_ARG;
    mk_status st = validate(ctx, p);
    if (st) return st;
    if (!host_rgb) return mk_fail(ctx, MK_ERR_ARG, "null output");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    size_t bytes = (size_t)p->width * p->height * 3;
    st = mk_ensure_buffer(ctx, &ctx->d_stage, &ctx->stage_cap, bytes);
    if (st) return st;
    st = mk_render_async(ctx, p, (uint8_t*)ctx->d_stage, ctx->stream);
    if (st) return st;
    MK_HIP(ctx, hipMemcpyAsync(host_rgb, ctx->d_stage, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}
