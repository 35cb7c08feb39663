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
    const uint8_t* occ;
    const float* mipb;
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

// Conservative [a_in, a_out] of accumulated lengths whose samples can pass
// the VOI test; returns false when none can.
__device__ inline bool ray_interval(const RenderArgs& A, float u, float v, float& a_in, float& a_out) {
    float p0[3];
    ray_pos(A, u, v, 0.0f, p0);
    float d[3] = {A.M[1] * (float)A.scale * A.mls[0], A.M[4] * (float)A.scale * A.mls[1],
                  A.M[7] * (float)A.scale * A.mls[2]};
    if (A.invert_z) d[2] = -d[2];
    const int lo_i[3] = {A.voi.left, A.voi.posterior, A.voi.head};
    const int hi_i[3] = {A.voi.right, A.voi.anterior, A.voi.foot};
    const int n[3] = {A.v.nx, A.v.ny, A.v.nz};
    float t0 = 0.0f, t1 = 1.7321f;
    const float eps = 1e-3f;
    for (int i = 0; i < 3; ++i) {
        // trunc(p*N) in [lo, hi]  =>  p in (lo-1, hi+1)/N
        float lo = ((float)lo_i[i] - 1.0f) / (float)n[i] - eps;
        float hi = ((float)hi_i[i] + 1.0f) / (float)n[i] + eps;
        if (fabsf(d[i]) < 1e-12f) {
            if (!(p0[i] >= lo && p0[i] <= hi)) return false;
        } else {
            float ta = (lo - p0[i]) / d[i], tb = (hi - p0[i]) / d[i];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
        }
    }
    const float m = 4.0f * A.l1;
    a_in = fmaxf(t0 - m, 0.0f);
    a_out = t1 + m;
    return a_in <= a_out;
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
// a zero gradient gives NaN, which fails every "> 0" test downstream.
__device__ inline float3 gradient(const RenderArgs& A, const float p[3]) {
    float3 N;
    N.x = mk_sample_hu(A.v, p[0] + A.sv[0], p[1], p[2]) - mk_sample_hu(A.v, p[0] - A.sv[0], p[1], p[2]);
    N.y = mk_sample_hu(A.v, p[0], p[1] + A.sv[1], p[2]) - mk_sample_hu(A.v, p[0], p[1] - A.sv[1], p[2]);
    N.z = mk_sample_hu(A.v, p[0], p[1], p[2] + A.sv[2]) - mk_sample_hu(A.v, p[0], p[1], p[2] - A.sv[2]);
    if (A.invert_z) N.z = -N.z;
    float r = rsqrtf(N.x * N.x + N.y * N.y + N.z * N.z);
    return make_float3(N.x * r, N.y * r, N.z * r);
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
    if (ray_interval(A, u, v, a_in, a_out)) {
        const float3 L = light_dir(A);
        const float L1 = A.l1, L4 = L1 / 4.0f, L8 = L1 / 8.0f;
        float step = L1, aPre = 0.f, aAcc = 0.f, aobj0 = 0.f;
        if (MASK)
            for (int k = 0; k < A.nslot; ++k) s.aobj[k * BS + tid] = 0.f;
        int k0 = (int)(a_in / L1) - 2;
        k0 = k0 < 0 ? 0 : (k0 >= A.nsteps ? A.nsteps - 1 : k0);
        float acc = A.steps[k0];
        for (int it = 0; it < A.max_iter && (double)acc < 1.732 && acc <= a_out; ++it) {
            float p[3];
            ray_pos(A, u, v, acc, p);
            int ix, iy, iz;
            if (!in_voi(A, p, ix, iy, iz)) { acc += step; continue; }
            const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1), cz = mk_clampi(iz, 0, A.v.nz - 1);
            float a = 0.f;
            float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
            int label = 0;
            if (TOUCH || A.occ[mk_brick_index(A.v, cx, cy, cz)]) {
                if (MASK) label = A.v.mask[mk_voxel_offset(A.v, cx, cy, cz)];
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
            if (a > 0.f && step != L1) a = 1.0f - powf(1.0f - a, step == L4 ? 0.25f : 0.125f);
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
    if (ray_interval(A, u, v, a_in, a_out)) {
        const float L1 = A.l1;
        int k0 = (int)(a_in / L1) - 2;
        k0 = k0 < 0 ? 0 : (k0 >= A.nsteps ? A.nsteps - 1 : k0);
        float acc = A.steps[k0];
        for (int it = 0; it < A.max_iter && (double)acc < 1.732 && acc <= a_out; ++it) {
            float p[3];
            ray_pos(A, u, v, acc, p);
            int ix, iy, iz;
            if (in_voi(A, p, ix, iy, iz)) {
                const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1),
                          cz = mk_clampi(iz, 0, A.v.nz - 1);
                if (TOUCH || A.mipb[mk_brick_index(A.v, cx, cy, cz)] > aAcc) {
                    int label = MASK ? A.v.mask[mk_voxel_offset(A.v, cx, cy, cz)] : 0;
                    if (TOUCH && MASK) touch_mask(A, cx, cy, cz);
                    const float4 aw = s.aw[label];
                    if (aw.x > 0.f) {
                        if (TOUCH) touch_footprint(A, p[0], p[1], p[2]);
                        float t = mk_sample_hu(A.v, p[0], p[1], p[2]);
                        t = (t - aw.z) / aw.y + 0.5f;
                        if (aAcc < t) aAcc = t;
                    }
                }
            }
            acc += L1;
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
    if (ray_interval(A, u, v, a_in, a_out)) {
        const float3 L = light_dir(A);
        const float L1 = A.l1;
        int k0 = (int)(a_in / L1) - 2;
        k0 = k0 < 0 ? 0 : (k0 >= A.nsteps ? A.nsteps - 1 : k0);
        float acc = A.steps[k0];
        for (int it = 0; it < A.max_iter && (double)acc < 1.732 && acc <= a_out; ++it) {
            float p[3];
            ray_pos(A, u, v, acc, p);
            int ix, iy, iz;
            if (in_voi(A, p, ix, iy, iz)) {
                const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1),
                          cz = mk_clampi(iz, 0, A.v.nz - 1);
                if (TOUCH || A.occ[mk_brick_index(A.v, cx, cy, cz)]) {
                    if (TOUCH) touch_footprint(A, p[0], p[1], p[2]);
                    float t = mk_sample_hu(A.v, p[0], p[1], p[2]);
                    if (t > 400.0f) {
                        if (TOUCH) touch_gradient(A, p);
                        const float3 N = gradient(A, p);
                        aAcc = N.x * L.x + N.y * L.y + N.z * L.z;
                        if (aAcc > 0.f) break;
                    }
                }
            }
            acc += L1;
        }
    }
    float4 c;
    if (aAcc <= 0.f) c = bg4(A);
    else {
        float g = (float)((double)aAcc + 0.05);
        c = make_float4(g, g, g, 1.0f);
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

mk_status ensure_steps(mk_context* ctx, float l1) {
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
    st = ensure_steps(ctx, l1);
    if (st) return st;

    RenderArgs A;
    memset(&A, 0, sizeof(A));
    A.v = ctx->view();
    A.lut = ctx->d_lut;
    A.tab = ctx->lut_tab;
    A.occ = ctx->d_occ;
    A.mipb = ctx->d_mipb;
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
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (s != ctx->stream) {
        // the skip tables were (re)built on the context stream
        hipEvent_t ev;
        MK_HIP(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        MK_HIP(ctx, hipEventRecord(ev, ctx->stream));
        MK_HIP(ctx, hipStreamWaitEvent(s, ev, 0));
        MK_HIP(ctx, hipEventDestroy(ev));
    }
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
    st = mk_render_async(ctx, p, (uint8_t*)ctx->d_stage, nullptr);
    if (st) return st;
    MK_HIP(ctx, hipMemcpyAsync(host_rgb, ctx->d_stage, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}
