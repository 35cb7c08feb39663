// mk_fit.hip -- differentiable transfer-function rendering and its adjoint.
//
// Fitting rule: the reference's front-to-back compositing (core/kernel.cu:49-81,
// 195-252) on its float step lattice, with the choices that make the image a
// smooth function of the transfer function:
//   * a fixed step (no getNextStep refinement: its lattice jumps with alpha),
//   * every sample with alpha > 0 is composited (no 0.0005 threshold),
//   * no early exit and no per-label caps (label 0 of the palette only),
//   * sample alpha = min(1, s * LUT_alpha) with a learnable opacity scale s,
//   * the same 0.35 / 0.8 / 0.16 headlight shading from the volume gradient,
//   * output = the float RGB sum (no saturation, no 8-bit rounding).
// For a ray with samples i (colour c_i, alpha a_i, shading k_i, transmittance
// T_i = prod_{j<i}(1 - a_j)): C = sum_i T_i a_i k_i c_i.rgb and
//   dC/dc_i.rgb = T_i a_i k_i,
//   dC/da_i     = T_i k_i c_i.rgb - (C - C_{<=i}) / (1 - a_i),
// where C_{<=i} is the colour composited up to and including sample i. The
// adjoint re-marches each ray once, un-compositing the remainder with that
// (1 - a) division, and scatters the LUT and scale gradients into
// per-workgroup LDS (double), then into HBM with double atomics. C comes from
// the forward pass, which can keep it in double (24 bytes per pixel); without
// it the adjoint first re-marches the ray for C. No sample is stored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "mk_context.h"
#include "mk_render_dev.h"

namespace {

struct FitArgs {
    MkVolView v;
    const float4* lut;  // n entries, device
    int n;
    float scale;        // opacity scale s
    const float* steps;
    int nsteps;
    int width, height;
    int invert_z;
    float mls[3];
    float sv[3];
    float xT, yT;
    double dscale;      // camera zoom
    float M[9];
    mk_voi voi;
    float wl, ww;       // label 0 window
    float l1;
    int max_iter;
};

// The sample position of the VR kernels' ray_pos: the same fused operations
// in the same order, so both rules place their samples identically.
__device__ inline void fit_pos(const FitArgs& A, float u, float v, float acc, float p[3]) {
    const float fy = (float)(((double)acc - 0.866) * A.dscale);
    const float x = fmaf(v, A.M[2], fmaf(fy, A.M[1], u * A.M[0]));
    const float y = fmaf(v, A.M[5], fmaf(fy, A.M[4], u * A.M[3]));
    const float z = fmaf(v, A.M[8], fmaf(fy, A.M[7], u * A.M[6]));
    p[0] = fmaf(x, A.mls[0], 0.5f);
    p[1] = fmaf(y, A.mls[1], 0.5f);
    p[2] = fmaf(z, A.mls[2], 0.5f);
    if (A.invert_z) p[2] = 1.0f - p[2];
}

__device__ inline bool fit_voi(const FitArgs& A, const float p[3]) {
    float fx = fminf(fmaxf(p[0] * (float)A.v.nx, -1e9f), 1e9f);
    float fy = fminf(fmaxf(p[1] * (float)A.v.ny, -1e9f), 1e9f);
    float fz = fminf(fmaxf(p[2] * (float)A.v.nz, -1e9f), 1e9f);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    return !(ix < A.voi.left || ix > A.voi.right || iy < A.voi.posterior || iy > A.voi.anterior || iz < A.voi.head ||
             iz > A.voi.foot);
}

// Shading factor 0.35 + (d > 0 ? 0.8 d + 0.16 d^8 : 0) of the normalized
// one-voxel central-difference gradient against the headlight. The six
// samples at p +- 1/N per axis share p's weights, so the differences are the
// interpolations (with p's weights) of voxel differences two apart: the
// sample's own eight corners plus 24 more voxels, loaded together; the common
// 32768/32767 scale cancels in the normalization.
__device__ inline float fit_shade(const FitArgs& A, const mkr::TriSample& tri, const mkr::GradTaps& taps, float3 L) {
    float N[3];
    mkr::grad_diff_finish(tri, taps, N);
    if (A.invert_z) N[2] = -N[2];
    float r = rsqrtf(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    float d = (N[0] * L.x + N[1] * L.y + N[2] * L.z) * r;
    float k = 0.35f;
    if (d > 0.f) {
        float d2 = d * d, d4 = d2 * d2;
        k += d * 0.8f + 0.16f * (d4 * d4);
    }
    return k;
}

struct Sample {
    int i0, i1;
    float f;      // weight of entry i1
    float4 c;     // LUT colour
    float a;      // alpha after scale and clamp
    bool grad_ok; // the clamp is inactive: 0 <= s*alpha < 1 (one-sided at 0)
};

// A sample from its fetched corners, already raised to -32767 at least;
// tri keeps them for fit_shade.
__device__ inline Sample fit_sample(const FitArgs& A, const float4* s_lut, const mkr::TriFetch& F, mkr::TriSample& tri) {
    Sample S;
    float t = mkr::sample_finish<false>(F, tri);
    t = (t - A.wl) / A.ww + 0.5f;
    t = t > 1.0f ? 1.0f : (t >= -1.0f ? t : -1.0f);
    mk_split(t, A.n, S.i0, S.i1, S.f);
    const float4 c0 = s_lut[S.i0], c1 = s_lut[S.i1];
    S.c = make_float4(c0.x + S.f * (c1.x - c0.x), c0.y + S.f * (c1.y - c0.y), c0.z + S.f * (c1.z - c0.z),
                      c0.w + S.f * (c1.w - c0.w));
    const float a = A.scale * S.c.w;
    S.grad_ok = a >= 0.0f && a < 1.0f;
    S.a = fminf(fmaxf(a, 0.0f), 1.0f);
    return S;
}

__device__ inline bool fit_ray(const FitArgs& A, int x, int y, float& u, float& v, int& k0, float& a_out) {
    u = ((float)x - (float)A.width / 2.0f - A.xT) / (float)A.height;
    v = ((float)y - (float)A.height / 2.0f - A.yT) / (float)A.height;
    float p0[3];
    fit_pos(A, u, v, 0.0f, p0);
    float d[3] = {A.M[1] * (float)A.dscale * A.mls[0], A.M[4] * (float)A.dscale * A.mls[1],
                  A.M[7] * (float)A.dscale * A.mls[2]};
    if (A.invert_z) d[2] = -d[2];
    const int lo_i[3] = {A.voi.left, A.voi.posterior, A.voi.head};
    const int hi_i[3] = {A.voi.right, A.voi.anterior, A.voi.foot};
    const int n[3] = {A.v.nx, A.v.ny, A.v.nz};
    float t0 = 0.0f, t1 = 1.7321f;
    for (int i = 0; i < 3; ++i) {
        float lo = ((float)lo_i[i] - 1.0f) / (float)n[i] - 1e-3f;
        float hi = ((float)hi_i[i] + 1.0f) / (float)n[i] + 1e-3f;
        if (fabsf(d[i]) < 1e-12f) {
            if (!(p0[i] >= lo && p0[i] <= hi)) return false;
        } else {
            float ta = (lo - p0[i]) / d[i], tb = (hi - p0[i]) / d[i];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
        }
    }
    t0 = fmaxf(t0 - 4.0f * A.l1, 0.0f);
    a_out = t1 + 4.0f * A.l1;
    if (t0 > a_out) return false;
    k0 = (int)(t0 / A.l1) - 2;
    k0 = k0 < 0 ? 0 : (k0 >= A.nsteps ? A.nsteps - 1 : k0);
    return true;
}

__device__ inline float3 fit_light(const FitArgs& A) {
    float3 l = make_float3(A.M[1], A.M[4], A.M[7]);
    float r = rsqrtf(l.x * l.x + l.y * l.y + l.z * l.z);
    return make_float3(l.x * r, l.y * r, l.z * r);
}

// A workgroup renders a 16x16 tile as four waves of 8x8 rays.
constexpr int kFitTile = 16;
constexpr int kFitWaves = 4;
__device__ inline void fit_tile(const FitArgs& A, int& x, int& y) {
    const uint32_t tiles_x = (A.width + kFitTile - 1) / kFitTile;
    const uint32_t t = mk_xcd_row_tile(blockIdx.x, tiles_x, (A.height + kFitTile - 1) / kFitTile);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    x = (int)(t % tiles_x) * kFitTile + (wave & 1) * 8 + (lane & 7);
    y = (int)(t / tiles_x) * kFitTile + (wave >> 1) * 8 + (lane >> 3);
}

// Orders this wave's LDS writes before its following reads (and its reads
// before the next writes): the wave's lanes share its staged box, and no
// other wave touches it.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// dynamic LDS of one fitting workgroup: the LUT (16 B per entry, up to 4096
// entries), in the adjoint its gradient sums when they fit beside it, and the
// staged blocks of the volume
constexpr size_t kFitLdsBytes = 64 * 1024;
constexpr int kFitChunk = 8;           // lattice points per staged chunk
constexpr int kFitBoxBlocks = 60;      // 4x4x4 blocks staged per chunk at most (7.5 KiB)

// The staged box: the 4x4x4 blocks [b0, b0 + nb) (block units) copied from
// the bricked volume in their own layout (128 B each, x-then-y-then-z), so a
// voxel is read at the sum of one term per axis, as in HBM.
struct FitBox {
    const uint8_t* base;
    int bx0, by0, bz0;
    uint32_t sy, sz;   // bytes per block row, per block slab
    // per-axis terms of the box offset for unclamped voxel indices from
    // 4 * b0 - 4 on (clamping folded in), built with the box
    const uint32_t* tx;
    const uint32_t* ty;
    const uint32_t* tz;
};
// table entries per axis for a box nb blocks long, and for any box within the
// cap (the sum over axes of nb is at most cap + 2)
__host__ __device__ constexpr int box_term_len(int nb) { return 4 * nb + 16; }
__host__ __device__ constexpr int box_terms_max(int cap) { return 4 * (cap + 2) + 48; }
__device__ inline uint32_t box_tx(const FitBox& B, int x) { return (uint32_t)(((x >> 2) - B.bx0) * 128 + (x & 3) * 2); }
__device__ inline uint32_t box_ty(const FitBox& B, int y) { return (uint32_t)((y >> 2) - B.by0) * B.sy + (uint32_t)((y & 3) * 8); }
__device__ inline uint32_t box_tz(const FitBox& B, int z) { return (uint32_t)((z >> 2) - B.bz0) * B.sz + (uint32_t)((z & 3) * 32); }
__device__ inline int box_vox(const FitBox& B, uint32_t off) { return *(const int16_t*)(B.base + off); }

// Fills the box's term tables (the wave's lanes; the caller orders them
// before use). Entry e of an axis holds the term of voxel index
// clamp(4 * b0 - 4 + e, 0, n - 1), held inside the box so that every entry
// addresses it.
__device__ inline void box_terms_build(const MkVolView& v, const FitBox& B, int nbx, int nby, int nbz, uint32_t* t) {
    const int lx = box_term_len(nbx), ly = box_term_len(nby), lz = box_term_len(nbz);
    for (int e = threadIdx.x & 63; e < lx + ly + lz; e += 64) {
        uint32_t w;
        if (e < lx) {
            const int i = mk_clampi(mk_clampi(4 * B.bx0 - 4 + e, 0, v.nx - 1), 4 * B.bx0, 4 * (B.bx0 + nbx) - 1);
            w = box_tx(B, i);
        } else if (e < lx + ly) {
            const int i = mk_clampi(mk_clampi(4 * B.by0 - 4 + (e - lx), 0, v.ny - 1), 4 * B.by0, 4 * (B.by0 + nby) - 1);
            w = box_ty(B, i);
        } else {
            const int i =
                mk_clampi(mk_clampi(4 * B.bz0 - 4 + (e - lx - ly), 0, v.nz - 1), 4 * B.bz0, 4 * (B.bz0 + nbz) - 1);
            w = box_tz(B, i);
        }
        t[e] = w;
    }
}

// sample_fetch and grad_taps_fetch from the box: the same clamped indices,
// so the same voxel values (raised to -32767 at least when the box was
// staged, as the samplers raise them). Floor indices lie in [-3, n + 2], so
// the entries read, ix + 3 .. ix + 6 - 4 * b0, lie in the tables.
__device__ inline void box_fetch(const MkVolView& v, const FitBox& B, const float p[3], mkr::TriFetch& F,
                                 mkr::GradTaps& G) {
    mkr::grad_axis_base(p[0], v.nx, F.ix, F.ax);
    mkr::grad_axis_base(p[1], v.ny, F.iy, F.ay);
    mkr::grad_axis_base(p[2], v.nz, F.iz, F.az);
    const uint32_t* tx = B.tx + (F.ix + 3 - 4 * B.bx0);
    const uint32_t* ty = B.ty + (F.iy + 3 - 4 * B.by0);
    const uint32_t* tz = B.tz + (F.iz + 3 - 4 * B.bz0);
    uint32_t OX[4], OY[4], OZ[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        OX[k] = tx[k];
        OY[k] = ty[k];
        OZ[k] = tz[k];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) F.r[k] = box_vox(B, OX[1 + (k & 1)] + OY[1 + ((k >> 1) & 1)] + OZ[1 + (k >> 2)]);
#define MK_BV(a, b, c) ((float)box_vox(B, OX[a] + OY[b] + OZ[c]))
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            G.x[k][j][0] = MK_BV(0, 1 + j, 1 + k);
            G.x[k][j][1] = MK_BV(3, 1 + j, 1 + k);
            G.y[k][j][0] = MK_BV(1 + j, 0, 1 + k);
            G.y[k][j][1] = MK_BV(1 + j, 3, 1 + k);
            G.z[k][j][0] = MK_BV(1 + j, 1 + k, 0);
            G.z[k][j][1] = MK_BV(1 + j, 1 + k, 3);
        }
#undef MK_BV
}

// Two int16 voxels raised to -32767 at least (the samplers read -32768 as
// -32767), with one packed max.
__device__ inline uint32_t vox_floor2(uint32_t w) {
    typedef short short2v __attribute__((ext_vector_type(2)));
    const short2v v = __builtin_elementwise_max(__builtin_bit_cast(short2v, w), (short2v){-32767, -32767});
    return __builtin_bit_cast(uint32_t, v);
}

// LDS bytes of one wave's region: the box, then its term tables.
__host__ __device__ constexpr int box_region(int cap) { return cap * 128 + 4 * box_terms_max(cap); }

__device__ inline int wave_min(int x) {
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o));
    return x;
}
__device__ inline int wave_max(int x) {
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
    return x;
}

// Visits each active ray's in-VOI lattice points in order: visit(S, k) gets
// the sample and its shading factor and returns false to end that ray's
// march. A ray's points are those of the loop
//   for (it = 0; it < max_iter && (double)acc < 1.732 && acc <= a_out; ++it)
//     { p = pos(acc); acc += L1; if in VOI: visit }
// from acc = steps[k0]; the step lattice is shared, so point it of the ray is
// lattice point k0 + it (acc = steps[k0 + it]). The wave's rays march the
// lattice together, kFitChunk points at a time: the box of 4x4x4 blocks that
// holds every corner and gradient tap of the chunk's points is copied to LDS
// with coalesced loads, and the samples read it there, each 64-lane gather of
// the HBM path (touching ~30 cache lines) becoming LDS reads. A chunk whose
// box exceeds the LDS budget reads HBM. s_box is this wave's own region;
// every lane of the wave calls this (active or not).
template <typename Visit>
__device__ inline void fit_march(const FitArgs& A, const float4* s_lut, uint8_t* s_box, int box_cap, bool active,
                                 float u, float v, int k0, float a_out, float3 L, Visit&& visit) {
    const int lane = threadIdx.x & 63;
    bool done = !active;
    int c = wave_min(done ? 0x7FFFFFFF : k0);
    while (__ballot(!done) != 0ull) {
        c = __builtin_amdgcn_readfirstlane(c);
        // this lane's points in [c, c + kFitChunk): k in [kb, ke], and whether
        // its march ends in this chunk
        int kb = 0x7FFFFFFF, ke = -1;
        bool ends = false;
        if (!done) {
#pragma unroll
            for (int m = 0; m < kFitChunk; ++m) {
                const int k = c + m;
                if (k < k0 || ends) continue;
                const bool ok = k - k0 < A.max_iter && k < A.nsteps && (double)A.steps[k] < 1.732 && A.steps[k] <= a_out;
                if (ok) {
                    kb = min(kb, k);
                    ke = k;
                } else {
                    ends = true;
                }
            }
        }
        // the box: positions move monotonically along a ray, so the first and
        // last points bound its indices; taps reach one below, two above
        int lo[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, hi[3] = {-1, -1, -1};
        if (ke >= 0) {
            float p0[3], p1[3];
            fit_pos(A, u, v, A.steps[kb], p0);
            fit_pos(A, u, v, A.steps[ke], p1);
            const int n[3] = {A.v.nx, A.v.ny, A.v.nz};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                int i0, i1;
                float w;
                mkr::grad_axis_base(p0[i], n[i], i0, w);
                mkr::grad_axis_base(p1[i], n[i], i1, w);
                lo[i] = mk_clampi(min(i0, i1) - 1, 0, n[i] - 1);
                hi[i] = mk_clampi(max(i0, i1) + 2, 0, n[i] - 1);
            }
        }
        FitBox B;
        B.base = s_box;
        B.bx0 = wave_min(lo[0]) >> 2;
        B.by0 = wave_min(lo[1]) >> 2;
        B.bz0 = wave_min(lo[2]) >> 2;
        const int nbx = (wave_max(hi[0]) >> 2) - B.bx0 + 1;
        const int nby = (wave_max(hi[1]) >> 2) - B.by0 + 1;
        const int nbz = (wave_max(hi[2]) >> 2) - B.bz0 + 1;
        const bool any = __ballot(ke >= 0) != 0ull;
        const bool staged = any && nbx * nby * nbz <= box_cap;
        if (staged) {
            B.sy = (uint32_t)nbx * 128u;
            B.sz = (uint32_t)(nbx * nby) * 128u;
            const int nblk = nbx * nby * nbz;
            uint32_t* t = (uint32_t*)(s_box + box_cap * 128);
            B.tx = t;
            B.ty = t + box_term_len(nbx);
            B.tz = B.ty + box_term_len(nby);
            box_terms_build(A.v, B, nbx, nby, nbz, t);
            for (int i = lane; i < nblk * 8; i += 64) {
                const int blk = i >> 3, part = i & 7;
                const int X = B.bx0 + blk % nbx, Y = B.by0 + (blk / nbx) % nby, Z = B.bz0 + blk / (nbx * nby);
                const uint32_t brick =
                    __umul24(__umul24((uint32_t)(Z >> 1), (uint32_t)A.v.nby) + (uint32_t)(Y >> 1), (uint32_t)A.v.nbx) +
                    (uint32_t)(X >> 1);
                const uint32_t off = (brick << 10) | ((uint32_t)(Z & 1) << 9) | ((uint32_t)(Y & 1) << 8) |
                                     ((uint32_t)(X & 1) << 7) | ((uint32_t)part << 4);
                const uint4 w = *(const uint4*)((const uint8_t*)A.v.vox + off);
                *(uint4*)(s_box + blk * 128 + part * 16) =
                    make_uint4(vox_floor2(w.x), vox_floor2(w.y), vox_floor2(w.z), vox_floor2(w.w));
            }
            wave_lds_sync();
        }
        if (ke >= 0) {
            for (int k = kb; k <= ke; ++k) {
                float P[3];
                fit_pos(A, u, v, A.steps[k], P);
                if (!fit_voi(A, P)) continue;
                mkr::TriFetch F;
                mkr::GradTaps G;
                if (staged) {
                    box_fetch(A.v, B, P, F, G);
                } else {
                    mkr::sample_fetch(A.v, P, F);
                    mkr::grad_taps_fetch(A.v, F.ix, F.iy, F.iz, G);
#pragma unroll
                    for (int q = 0; q < 8; ++q) F.r[q] = max(F.r[q], -32767);
                }
                mkr::TriSample tri;
                const Sample S = fit_sample(A, s_lut, F, tri);
                if (!visit(S, fit_shade(A, tri, G, L))) {
                    ends = true;
                    break;
                }
            }
        }
        if (ends) done = true;
        if (staged) wave_lds_sync();   // the next chunk's box overwrites this one
        c += kFitChunk;
    }
}

// Forward: float RGB per pixel (width*height*3), and the same colours in
// double into out64 unless it is null.
// box_cap: 4x4x4 blocks of the volume each wave's region of LDS (after the
// LUT) holds.
__global__ void __launch_bounds__(256) k_fit_forward(FitArgs A, float* __restrict__ out, double* __restrict__ out64,
                                                     int box_cap) {
    extern __shared__ float4 s_lut[];
    for (int i = threadIdx.x; i < A.n; i += blockDim.x) s_lut[i] = A.lut[i];
    __syncthreads();
    uint8_t* s_box = (uint8_t*)(s_lut + A.n) + (threadIdx.x >> 6) * box_region(box_cap);
    int x, y;
    fit_tile(A, x, y);
    double C[3] = {0.0, 0.0, 0.0}, T = 1.0;
    float u = 0.f, v = 0.f, a_out = 0.f;
    int k0 = 0;
    const bool inside = x < A.width && y < A.height;
    const bool active = inside && fit_ray(A, x, y, u, v, k0, a_out);
    fit_march(A, s_lut, s_box, box_cap, active, u, v, k0, a_out, fit_light(A), [&](const Sample& S, float k) {
        if (S.a > 0.0f) {
            const double w = T * (double)S.a * (double)k;
            C[0] += w * S.c.x;
            C[1] += w * S.c.y;
            C[2] += w * S.c.z;
            T *= 1.0 - (double)S.a;
        }
        return true;
    });
    if (!inside) return;
    float* o = out + 3 * ((size_t)y * A.width + x);
    o[0] = (float)C[0];
    o[1] = (float)C[1];
    o[2] = (float)C[2];
    if (out64) {
        double* o64 = out64 + 3 * ((size_t)y * A.width + x);
        o64[0] = C[0];
        o64[1] = C[1];
        o64[2] = C[2];
    }
}

// Adjoint: given g = dL/dC per pixel, accumulate dL/dLUT (n x 4) and dL/ds.
// color64 holds the forward pass's double colours, or is null.
// LDS_GRAD: the workgroup sums its LUT and scale gradients in LDS (4n + 1
// doubles after the LUT) and adds the nonzero sums to HBM at the end. Without
// it (LUTs too long for both tables in LDS) the lanes add to HBM directly.
template <bool LDS_GRAD>
__global__ void __launch_bounds__(256) k_fit_backward(FitArgs A, const float* __restrict__ grad_img,
                                                     const double* __restrict__ color64,
                                                     double* __restrict__ grad_lut, double* __restrict__ grad_scale,
                                                     int box_cap) {
    extern __shared__ float4 smem[];
    float4* s_lut = smem;
    double* s_g = LDS_GRAD ? (double*)(smem + A.n) : grad_lut;  // 4*n LUT entries, then the scale
    double* s_scale = LDS_GRAD ? s_g + 4 * A.n : grad_scale;
    uint8_t* s_box = (LDS_GRAD ? (uint8_t*)(s_g + 4 * A.n + 2) : (uint8_t*)(smem + A.n)) +
                     (threadIdx.x >> 6) * box_region(box_cap);
    for (int i = threadIdx.x; i < A.n; i += blockDim.x) s_lut[i] = A.lut[i];
    if (LDS_GRAD)
        for (int i = threadIdx.x; i < 4 * A.n + 1; i += blockDim.x) s_g[i] = 0.0;
    __syncthreads();
    int x, y;
    fit_tile(A, x, y);
    float u = 0.f, v = 0.f, a_out = 0.f;
    int k0 = 0;
    double g[3] = {0.0, 0.0, 0.0};
    bool active = x < A.width && y < A.height && fit_ray(A, x, y, u, v, k0, a_out);
    if (active) {
        const float* gp = grad_img + 3 * ((size_t)y * A.width + x);
        g[0] = gp[0];
        g[1] = gp[1];
        g[2] = gp[2];
        active = g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0;
    }
    const float3 L = fit_light(A);
    // the final colour: kept by the forward pass, else a first march
    double C[3] = {0.0, 0.0, 0.0}, T = 1.0;
    if (color64) {
        if (active) {
            const double* cp = color64 + 3 * ((size_t)y * A.width + x);
            C[0] = cp[0];
            C[1] = cp[1];
            C[2] = cp[2];
        }
    } else {
        fit_march(A, s_lut, s_box, box_cap, active, u, v, k0, a_out, L, [&](const Sample& S, float k) {
            if (S.a > 0.0f) {
                const double w = T * (double)S.a * (double)k;
                C[0] += w * S.c.x;
                C[1] += w * S.c.y;
                C[2] += w * S.c.z;
                T *= 1.0 - (double)S.a;
            }
            return true;
        });
    }
    // gradients, un-compositing the remainder of the ray
    double Cp[3] = {0.0, 0.0, 0.0};
    T = 1.0;
    double gs = 0.0;
    // LUT gradients of consecutive samples that interpolate the same two
    // entries are summed in registers and added to LDS when the pair changes
    int p0 = -1, p1 = -1;
    double q0[4] = {0.0, 0.0, 0.0, 0.0}, q1[4] = {0.0, 0.0, 0.0, 0.0};
    auto flush = [&]() {
        if (p0 < 0) return;
        for (int ch = 0; ch < 4; ++ch) {
            if (q0[ch] != 0.0) atomicAdd(&s_g[4 * p0 + ch], q0[ch]);
            if (q1[ch] != 0.0) atomicAdd(&s_g[4 * p1 + ch], q1[ch]);
            q0[ch] = 0.0;
            q1[ch] = 0.0;
        }
    };
    fit_march(A, s_lut, s_box, box_cap, active, u, v, k0, a_out, L, [&](const Sample& S, float kf) {
        // samples with zero alpha still move the image when their LUT
        // entries grow: da > 0 from 0 is a smooth onset
        const double k = (double)kf;
        const double a = S.a;
        const double w = T * a * k;
        const double c[3] = {S.c.x, S.c.y, S.c.z};
        for (int ch = 0; ch < 3; ++ch) Cp[ch] += w * c[ch];
        // dL/dc.rgb
        const double drgb[3] = {g[0] * w, g[1] * w, g[2] * w};
        // dL/da
        // an opaque sample (alpha clamped at 1) hides the rest and passes no
        // gradient to its alpha
        double da = 0.0;
        if (S.grad_ok) {
            // one double division per sample (1 - a > 0 here), not one per channel
            const double r = 1.0 / (1.0 - a);
            for (int ch = 0; ch < 3; ++ch) da += g[ch] * (T * k * c[ch] - (C[ch] - Cp[ch]) * r);
        }
        T *= 1.0 - a;
        const double dw = da;
        // alpha = s * c.w (unclamped region): dL/dc.w = s dL/da, dL/ds = c.w dL/da
        const double dcw = dw * (double)A.scale;
        gs += dw * (double)S.c.w;
        const double f = S.f;
        const double d4[4] = {drgb[0], drgb[1], drgb[2], dcw};
        if (S.i0 != p0 || S.i1 != p1) {
            // a held entry that the new pair also uses stays in its register
            // (along a ray the pair mostly moves by one entry); the others
            // go to LDS
            double c0[4] = {0.0, 0.0, 0.0, 0.0}, c1[4] = {0.0, 0.0, 0.0, 0.0};
            if (p0 >= 0) {
                for (int ch = 0; ch < 4; ++ch) {
                    if (p0 == S.i0) c0[ch] += q0[ch];
                    else if (p0 == S.i1) c1[ch] += q0[ch];
                    else if (q0[ch] != 0.0) atomicAdd(&s_g[4 * p0 + ch], q0[ch]);
                    if (p1 == S.i0) c0[ch] += q1[ch];
                    else if (p1 == S.i1) c1[ch] += q1[ch];
                    else if (q1[ch] != 0.0) atomicAdd(&s_g[4 * p1 + ch], q1[ch]);
                }
            }
            for (int ch = 0; ch < 4; ++ch) {
                q0[ch] = c0[ch];
                q1[ch] = c1[ch];
            }
            p0 = S.i0;
            p1 = S.i1;
        }
        for (int ch = 0; ch < 4; ++ch) {
            q0[ch] += (1.0 - f) * d4[ch];
            q1[ch] += f * d4[ch];
        }
        return T != 0.0;
    });
    flush();
    if (gs != 0.0) atomicAdd(s_scale, gs);
    if (!LDS_GRAD) return;
    __syncthreads();
    for (int i = threadIdx.x; i < 4 * A.n; i += blockDim.x)
        if (s_g[i] != 0.0) atomicAdd(&grad_lut[i], s_g[i]);
    if (threadIdx.x == 0 && s_g[4 * A.n] != 0.0) atomicAdd(grad_scale, s_g[4 * A.n]);
}

// Blocks of each wave's staged box that fit, with their term tables, in the
// LDS left after `used` bytes.
int box_blocks(size_t used) {
    const size_t left = used < kFitLdsBytes ? kFitLdsBytes - used : 0;
    int cap = kFitBoxBlocks;
    while (cap > 0 && (size_t)box_region(cap) * kFitWaves > left) --cap;
    return cap;
}

mk_status fit_prepare(mk_context* ctx, const mk_render_params* p, const float* dev_lut, int n, float scale, FitArgs& A) {
    if (!p) return mk_fail(ctx, MK_ERR_ARG, "render params are null");
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "fit before any volume was uploaded");
    if (!dev_lut || n < 2 || n > 4096) return mk_fail(ctx, MK_ERR_ARG, "fit LUT must have 2..4096 entries (got %d)", n);
    if (p->width <= 0 || p->height <= 0 || p->width > 16384 || p->height > 16384)
        return mk_fail(ctx, MK_ERR_ARG, "bad image size %dx%d", p->width, p->height);
    if (!(p->step_scale > 0.0f) || !std::isfinite(p->step_scale) || ctx->nz / p->step_scale > 1e5f)
        return mk_fail(ctx, MK_ERR_ARG, "bad step_scale %g", (double)p->step_scale);
    if (!std::isfinite(scale) || scale < 0.0f) return mk_fail(ctx, MK_ERR_ARG, "bad opacity scale %g", (double)scale);
    if (!(p->alpha_wwwl[0].ww != 0.0f)) return mk_fail(ctx, MK_ERR_ARG, "label 0 window width is zero");
    const float l1 = 1.0f / (float)ctx->nz * p->step_scale;
    mk_status st = mk_ensure_steps(ctx, l1);
    if (st) return st;
    A.v = ctx->view();
    A.lut = (const float4*)dev_lut;
    A.n = n;
    A.scale = scale;
    A.steps = ctx->d_steps;
    A.nsteps = ctx->steps_n;
    A.width = p->width;
    A.height = p->height;
    A.invert_z = p->invert_z != 0;
    for (int i = 0; i < 3; ++i) {
        A.mls[i] = p->max_len_spacing[i];
        A.sv[i] = p->spacing_voxel[i];
    }
    A.xT = p->x_translate;
    A.yT = p->y_translate;
    A.dscale = (double)p->scale;
    memcpy(A.M, p->transform, sizeof(A.M));
    A.voi = p->voi;
    A.wl = p->alpha_wwwl[0].wl;
    A.ww = p->alpha_wwwl[0].ww;
    A.l1 = l1;
    A.max_iter = ctx->steps_n + 64;
    return MK_OK;
}

}  // namespace

extern "C" mk_status mk_tf_forward(mk_context* ctx, const mk_render_params* p, const float* dev_lut, int n,
                                   float opacity_scale, float* dev_rgb, double* dev_rgb64, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    FitArgs A;
    mk_status st = fit_prepare(ctx, p, dev_lut, n, opacity_scale, A);
    if (st) return st;
    if (!dev_rgb) return mk_fail(ctx, MK_ERR_ARG, "null output");
    hipStream_t s = (hipStream_t)stream;
    st = mk_order_after_context(ctx, s);
    if (st) return st;
    const uint32_t tiles = (uint32_t)(((p->width + kFitTile - 1) / kFitTile) * ((p->height + kFitTile - 1) / kFitTile));
    const size_t lds_lut = (size_t)n * sizeof(float4);
    const int box_cap = box_blocks(lds_lut);
    hipLaunchKernelGGL(k_fit_forward, dim3(tiles), dim3(256), lds_lut + (size_t)box_region(box_cap) * kFitWaves, s, A,
                       dev_rgb, dev_rgb64, box_cap);
    MK_HIP(ctx, hipGetLastError());
    return mk_note_reader(ctx, s);
}

extern "C" mk_status mk_tf_backward(mk_context* ctx, const mk_render_params* p, const float* dev_lut, int n,
                                    float opacity_scale, const float* dev_grad_rgb, const double* dev_rgb64,
                                    double* dev_grad_lut, double* dev_grad_scale, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    FitArgs A;
    mk_status st = fit_prepare(ctx, p, dev_lut, n, opacity_scale, A);
    if (st) return st;
    if (!dev_grad_rgb || !dev_grad_lut || !dev_grad_scale) return mk_fail(ctx, MK_ERR_ARG, "null gradient buffer");
    hipStream_t s = (hipStream_t)stream;
    st = mk_order_after_context(ctx, s);
    if (st) return st;
    const uint32_t tiles = (uint32_t)(((p->width + kFitTile - 1) / kFitTile) * ((p->height + kFitTile - 1) / kFitTile));
    // gradient sums in LDS beside the LUT when both fit (n <= 1365), else in
    // HBM; the volume's staged blocks take what is left
    const size_t lds_lut = (size_t)n * sizeof(float4);
    const size_t lds_all = lds_lut + (4 * (size_t)n + 2) * sizeof(double);
    if (lds_all <= kFitLdsBytes) {
        const int box_cap = box_blocks(lds_all);
        hipLaunchKernelGGL(k_fit_backward<true>, dim3(tiles), dim3(256), lds_all + (size_t)box_region(box_cap) * kFitWaves,
                           s, A, dev_grad_rgb, dev_rgb64, dev_grad_lut, dev_grad_scale, box_cap);
    } else {
        const int box_cap = box_blocks(lds_lut);
        hipLaunchKernelGGL(k_fit_backward<false>, dim3(tiles), dim3(256), lds_lut + (size_t)box_region(box_cap) * kFitWaves,
                           s, A, dev_grad_rgb, dev_rgb64, dev_grad_lut, dev_grad_scale, box_cap);
    }
    MK_HIP(ctx, hipGetLastError());
    return mk_note_reader(ctx, s);
}
