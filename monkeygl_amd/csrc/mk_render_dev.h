// mk_render_dev.h -- device code of VR / MIP / Surface ray casting, shared
// by the per-frame kernels (mk_render.hip) and the batched serving kernel
// (mk_batch.hip). See mk_render.hip for the design notes.
#pragma once

#include <hip/hip_runtime.h>

#include "mk_context.h"
#include "mk_lattice.h"

namespace mkr {

// VR rays still marching after the first pass's trip budget, queued with the
// state the reference's loop carries from one iteration to the next
// (core/kernel.cu:133-264: position t, step, alpha of the previous sample,
// accumulated alpha and colour, per-label accumulated alpha) plus the trip
// count. Structure of arrays over the queue; cap = 0 disables the queue.
struct VrCont {
    uint32_t* counters;   // [0] rays queued from the front, [1] rays claimed by
                          // the second pass, [2] rays queued from the back,
                          // [3] first-pass waves finished
    int* pix;             // y * width + x
    float* acc;
    float* step;
    float* apre;
    float* aacc;
    int* it;
    float4* sum;
    float* aobj;          // [slot * cap + ray]
    int cap;
    int budget;           // loop trips of the first pass
    uint32_t tail_waves;  // first-pass waves finished after which running ones hand off
    float long_steps;     // rays with at least this many L1 steps left to the
                          // end of their interval queue from the front
};

struct RenderArgs {
    MkVolView v;
    const float4* lut;
    MkLutTable tab;
    const uint8_t* occ;    // per 8^3 brick (VR, Surface)
    const float* mipb;     // per 8^3 brick (MIP)
    const uint8_t* occ2;   // per 32^3 macro cell (4x4x4 bricks)
    const uint8_t* blabel; // per brick: its only label, or 0xFF when mixed
    const uint16_t* bskip; // per brick: label, occ and macro occ2 (MK_BSKIP_*)
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
    float rl1;   // 1 / l1 (an estimate for lattice_skip)
    float exit_alpha;
    int max_iter;      // hard bound on loop trips per ray
    VrCont cont;       // VR second pass (cap 0 in every other use)
};

struct AwTable {
    float alpha[MK_NUM_LABELS];
    float ww[MK_NUM_LABELS];
    float wl[MK_NUM_LABELS];
};

// The reference's loop bound t < 1.732 (a double comparison of the float t):
// 1.732f is the largest float below 1.732, so the bound is t <= 1.732f.
constexpr float kAccEnd = 1.732f;

// Sample position at accumulated length acc (core/kernel.cu:197-210). The
// multiply-adds are written as fused operations in a fixed order (the
// reference's dot product left to right, then pos * maxLenSpacing + 0.5), so
// the float positions are defined to the bit: in views along an axis they
// fall exactly on voxel faces, where their last bit decides the VOI test and
// the mask label of a whole column of samples.
__device__ inline void ray_pos(const RenderArgs& A, float u, float v, float acc, float p[3]) {
    const float fy = (float)(((double)acc - 0.866) * A.scale);
    const float x = fmaf(v, A.M[2], fmaf(fy, A.M[1], u * A.M[0]));
    const float y = fmaf(v, A.M[5], fmaf(fy, A.M[4], u * A.M[3]));
    const float z = fmaf(v, A.M[8], fmaf(fy, A.M[7], u * A.M[6]));
    p[0] = fmaf(x, A.mls[0], 0.5f);
    p[1] = fmaf(y, A.mls[1], 0.5f);
    p[2] = fmaf(z, A.mls[2], 0.5f);
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
// over the three axes, with the cell shrunk by a margin (1e-5 of the volume's
// extent) that covers the rounding of the sample positions. Quotients are
// taken as products with hardware reciprocals: their error, a few ulp of a
// distance below 2, stays ~25x inside the margin.
__device__ inline float cell_exit(const RenderArgs& A, const RayGeom& g, int cx, int cy, int cz, int shift) {
    const int c[3] = {cx >> shift, cy >> shift, cz >> shift};
    const float n[3] = {(float)A.v.nx, (float)A.v.ny, (float)A.v.nz};
    const float size = (float)(1 << shift);
    const float margin = 1e-5f;
    float t = INFINITY;
    for (int i = 0; i < 3; ++i) {
        if (g.d[i] > 1e-12f) {
            float face = ((float)(c[i] + 1) * size) * __builtin_amdgcn_rcpf(n[i]) - margin;
            t = fminf(t, (face - g.p0[i]) * __builtin_amdgcn_rcpf(g.d[i]));
        } else if (g.d[i] < -1e-12f) {
            float face = ((float)c[i] * size) * __builtin_amdgcn_rcpf(n[i]) + margin;
            t = fminf(t, (face - g.p0[i]) * __builtin_amdgcn_rcpf(g.d[i]));
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
    return __umul24(__umul24((uint32_t)(cz >> 5), (uint32_t)A.nmy) + (uint32_t)(cy >> 5), (uint32_t)A.nmx) +
           (uint32_t)(cx >> 5);
}

// tex1D linear/clamp at normalized s of the LUT of n entries at off.
template <bool LDS>
__device__ inline float4 lut_sample(const RenderArgs& A, const float4* s_lut, int off, int n, float s) {
    if (n <= 0) return make_float4(0.f, 0.f, 0.f, 0.f);
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
// The six samples at p +- sv (sv = 1/N per axis) sit one texel from p, so they
// share p's fractional weights: each difference is the trilinear
// interpolation, with p's weights, of voxel differences two apart. That reads
// the 32 distinct voxels of the 4x2x2, 2x4x2 and 2x2x4 neighbourhoods once
// instead of 48 voxels through six separate samples. Brick offsets are
// separable per axis (brick index * 512 plus the in-brick position), so the
// 32 addresses are sums of 12 per-axis terms; clamping is per index, as the
// texture's clamp addressing is. The common 32768/32767 scale cancels in the
// normalization and is not applied.
__device__ inline void grad_axis_base(float f, int n, int& i, float& a) {
    float xt = f * (float)n - 0.5f;
    float fl = floorf(xt);
    a = xt - fl;
    fl = fminf(fmaxf(fl, -3.0f), (float)(n + 2));
    i = (int)fl;
}

// A trilinear sample that keeps what the gradient at the same point reuses:
// the unclamped floor index and weight per axis and the eight corner voxels
// (the gradient's centre block). Same arithmetic as mk_sample_hu.
struct TriSample {
    int ix, iy, iz;
    float ax, ay, az;
    float c[2][2][2];   // [z][y][x], raw voxels (-32768 read as -32767)
};

// The eight corner voxels of a sample, loaded but not yet used: a caller can
// issue other loads before the first use, so that they share a round trip.
struct TriFetch {
    int ix, iy, iz;
    float ax, ay, az;
    int r[8];   // raw voxels, [z][y][x] order
};

__device__ inline void sample_fetch(const MkVolView& v, const float p[3], TriFetch& F) {
    grad_axis_base(p[0], v.nx, F.ix, F.ax);
    grad_axis_base(p[1], v.ny, F.iy, F.ay);
    grad_axis_base(p[2], v.nz, F.iz, F.az);
    const uint32_t X0 = mk_term_x(mk_clampi(F.ix, 0, v.nx - 1)), X1 = mk_term_x(mk_clampi(F.ix + 1, 0, v.nx - 1));
    const uint32_t Y0 = mk_term_y(v, mk_clampi(F.iy, 0, v.ny - 1)), Y1 = mk_term_y(v, mk_clampi(F.iy + 1, 0, v.ny - 1));
    const uint32_t Z0 = mk_term_z(v, mk_clampi(F.iz, 0, v.nz - 1)), Z1 = mk_term_z(v, mk_clampi(F.iz + 1, 0, v.nz - 1));
    F.r[0] = mk_vox_at(v, X0 + Y0 + Z0);
    F.r[1] = mk_vox_at(v, X1 + Y0 + Z0);
    F.r[2] = mk_vox_at(v, X0 + Y1 + Z0);
    F.r[3] = mk_vox_at(v, X1 + Y1 + Z0);
    F.r[4] = mk_vox_at(v, X0 + Y0 + Z1);
    F.r[5] = mk_vox_at(v, X1 + Y0 + Z1);
    F.r[6] = mk_vox_at(v, X0 + Y1 + Z1);
    F.r[7] = mk_vox_at(v, X1 + Y1 + Z1);
}

// The trilinear value of a fetched sample, keeping what the gradient reuses
// in T (same arithmetic as mk_sample_hu). CLAMP = false: the fetch's voxels
// are already at least -32767.
template <bool CLAMP = true>
__device__ inline float sample_finish(const TriFetch& F, TriSample& T) {
    T.ix = F.ix;
    T.iy = F.iy;
    T.iz = F.iz;
    T.ax = F.ax;
    T.ay = F.ay;
    T.az = F.az;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int q = F.r[k];
        T.c[k >> 2][(k >> 1) & 1][k & 1] = (float)(CLAMP && q < -32767 ? -32767 : q);
    }
    const float c00 = T.c[0][0][0] + T.ax * (T.c[0][0][1] - T.c[0][0][0]);
    const float c10 = T.c[0][1][0] + T.ax * (T.c[0][1][1] - T.c[0][1][0]);
    const float c01 = T.c[1][0][0] + T.ax * (T.c[1][0][1] - T.c[1][0][0]);
    const float c11 = T.c[1][1][0] + T.ax * (T.c[1][1][1] - T.c[1][1][0]);
    const float c0 = c00 + T.ay * (c10 - c00);
    const float c1 = c01 + T.ay * (c11 - c01);
    const float c = c0 + T.az * (c1 - c0);
    return c * (32768.0f / 32767.0f);
}

__device__ inline float sample_keep(const MkVolView& v, const float p[3], TriSample& T) {
    TriFetch F;
    sample_fetch(v, p, F);
    return sample_finish(F, T);
}

// The per-axis byte-offset terms of clamped indices (mk_term_x / _y / _z of
// clamp(i, 0, n - 1)) as tables in LDS, for every index a sample or its
// gradient taps can reach: grad_axis_base limits a sample's floor index to
// [-3, n + 2], its corners use i and i + 1, its taps i - 1 .. i + 2, so an axis
// holds indices -4 .. n + 4. A sample then reads each axis's two terms with
// one LDS instruction instead of composing them from bit fields (~10 VALU
// instructions a term); the values, and so the addresses, are the same.
constexpr int kTermPad = 4;
__host__ __device__ inline int axis_term_len(int n) { return n + 2 * kTermPad + 1; }

struct AxisTerms {
    const uint32_t* x;   // x[i] for i in [-kTermPad, nx + kTermPad]
    const uint32_t* y;
    const uint32_t* z;
};

// Fills the tables (all threads of the workgroup; the caller synchronizes).
__device__ inline AxisTerms stage_axis_terms(const MkVolView& v, uint32_t* t) {
    const int lx = axis_term_len(v.nx), ly = axis_term_len(v.ny), lz = axis_term_len(v.nz);
    for (int i = threadIdx.x; i < lx + ly + lz; i += blockDim.x) {
        uint32_t w;
        if (i < lx) w = mk_term_x(mk_clampi(i - kTermPad, 0, v.nx - 1));
        else if (i < lx + ly) w = mk_term_y(v, mk_clampi(i - lx - kTermPad, 0, v.ny - 1));
        else w = mk_term_z(v, mk_clampi(i - lx - ly - kTermPad, 0, v.nz - 1));
        t[i] = w;
    }
    return AxisTerms{t + kTermPad, t + lx + kTermPad, t + lx + ly + kTermPad};
}

// sample_fetch with the terms read from the tables.
__device__ inline void sample_fetch(const MkVolView& v, const AxisTerms& A, const float p[3], TriFetch& F) {
    grad_axis_base(p[0], v.nx, F.ix, F.ax);
    grad_axis_base(p[1], v.ny, F.iy, F.ay);
    grad_axis_base(p[2], v.nz, F.iz, F.az);
    const uint32_t X0 = A.x[F.ix], X1 = A.x[F.ix + 1];
    const uint32_t Y0 = A.y[F.iy], Y1 = A.y[F.iy + 1];
    const uint32_t Z0 = A.z[F.iz], Z1 = A.z[F.iz + 1];
    F.r[0] = mk_vox_at(v, X0 + Y0 + Z0);
    F.r[1] = mk_vox_at(v, X1 + Y0 + Z0);
    F.r[2] = mk_vox_at(v, X0 + Y1 + Z0);
    F.r[3] = mk_vox_at(v, X1 + Y1 + Z0);
    F.r[4] = mk_vox_at(v, X0 + Y0 + Z1);
    F.r[5] = mk_vox_at(v, X1 + Y0 + Z1);
    F.r[6] = mk_vox_at(v, X0 + Y1 + Z1);
    F.r[7] = mk_vox_at(v, X1 + Y1 + Z1);
}

// sample_keep from a fetch through the tables (TAB) or by composing the terms.
template <bool TAB>
__device__ inline float sample_keep(const MkVolView& v, const AxisTerms& A, const float p[3], TriSample& T) {
    TriFetch F;
    if constexpr (TAB) sample_fetch(v, A, p, F);
    else sample_fetch(v, p, F);
    return sample_finish(F, T);
}

// The 24 voxels around a sample's centre block that its central differences
// read (the outer layers at index 0 and 3 of each axis), loaded but not yet
// used, so that a caller can put them in the same round trip as other loads.
// Their addresses depend on the sample's position only. Voxels as floats
// (-32768 read as -32767), [k][j][end] per axis as grad_diff_finish reads them.
struct GradTaps {
    float x[2][2][2], y[2][2][2], z[2][2][2];
};

__device__ inline void grad_taps_fetch(const MkVolView& v, int ix, int iy, int iz, GradTaps& G) {
    uint32_t OX[4], OY[4], OZ[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        OX[k] = mk_term_x(mk_clampi(ix - 1 + k, 0, v.nx - 1));
        OY[k] = mk_term_y(v, mk_clampi(iy - 1 + k, 0, v.ny - 1));
        OZ[k] = mk_term_z(v, mk_clampi(iz - 1 + k, 0, v.nz - 1));
    }
#define MK_V(a, b, c) mk_rawo(v, OX[a] + OY[b] + OZ[c])
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            G.x[k][j][0] = MK_V(0, 1 + j, 1 + k);
            G.x[k][j][1] = MK_V(3, 1 + j, 1 + k);
            G.y[k][j][0] = MK_V(1 + j, 0, 1 + k);
            G.y[k][j][1] = MK_V(1 + j, 3, 1 + k);
            G.z[k][j][0] = MK_V(1 + j, 1 + k, 0);
            G.z[k][j][1] = MK_V(1 + j, 1 + k, 3);
        }
#undef MK_V
}

// grad_taps_fetch with the terms read from the tables.
__device__ inline void grad_taps_fetch(const MkVolView& v, const AxisTerms& A, int ix, int iy, int iz, GradTaps& G) {
    uint32_t OX[4], OY[4], OZ[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        OX[k] = A.x[ix - 1 + k];
        OY[k] = A.y[iy - 1 + k];
        OZ[k] = A.z[iz - 1 + k];
    }
#define MK_V(a, b, c) mk_rawo(v, OX[a] + OY[b] + OZ[c])
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            G.x[k][j][0] = MK_V(0, 1 + j, 1 + k);
            G.x[k][j][1] = MK_V(3, 1 + j, 1 + k);
            G.y[k][j][0] = MK_V(1 + j, 0, 1 + k);
            G.y[k][j][1] = MK_V(1 + j, 3, 1 + k);
            G.z[k][j][0] = MK_V(1 + j, 1 + k, 0);
            G.z[k][j][1] = MK_V(1 + j, 1 + k, 3);
        }
#undef MK_V
}

// The central differences at a kept sample from its fetched taps,
// unnormalized and without the common 32768/32767 scale.
__device__ inline void grad_diff_finish(const TriSample& T, const GradTaps& G, float N[3]) {
    const float ax = T.ax, ay = T.ay, az = T.az;
    const float(&c)[2][2][2] = T.c;
    // x: lerp(V2 - V0, V3 - V1, ax) per (y, z), then bilinear in (ay, az)
    float dx[2][2], dy[2][2], dz[2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float lo = c[k][j][1] - G.x[k][j][0], hi = G.x[k][j][1] - c[k][j][0];
            dx[k][j] = lo + ax * (hi - lo);
            // y: index j of the y taps runs over x (1 + j), k over z (1 + k)
            const float loy = c[k][1][j] - G.y[k][j][0], hiy = G.y[k][j][1] - c[k][0][j];
            dy[k][j] = loy + ay * (hiy - loy);
            // z: the z taps [k][j] are at x = 1 + j, y = 1 + k
            const float loz = c[1][k][j] - G.z[k][j][0], hiz = G.z[k][j][1] - c[0][k][j];
            dz[k][j] = loz + az * (hiz - loz);
        }
    {   // dx[k][j]: j over y, k over z
        const float a0 = dx[0][0] + ay * (dx[0][1] - dx[0][0]);
        const float a1 = dx[1][0] + ay * (dx[1][1] - dx[1][0]);
        N[0] = a0 + az * (a1 - a0);
    }
    {   // dy[k][j]: j over x, k over z
        const float a0 = dy[0][0] + ax * (dy[0][1] - dy[0][0]);
        const float a1 = dy[1][0] + ax * (dy[1][1] - dy[1][0]);
        N[1] = a0 + az * (a1 - a0);
    }
    {   // dz[k][j]: j over x, k over y
        const float a0 = dz[0][0] + ax * (dz[0][1] - dz[0][0]);
        const float a1 = dz[1][0] + ax * (dz[1][1] - dz[1][0]);
        N[2] = a0 + ay * (a1 - a0);
    }
}

// The central differences at a kept sample, unnormalized and without the
// common 32768/32767 scale: 24 more voxels around its centre block.
__device__ inline void grad_diff_kept(const MkVolView& v, const TriSample& T, float N[3]) {
    GradTaps G;
    grad_taps_fetch(v, T.ix, T.iy, T.iz, G);
    grad_diff_finish(T, G, N);
}

// The gradient from a kept sample (normalized, z flipped with the volume);
// TAB: tap addresses from the tables.
template <bool TAB = false>
__device__ inline float3 gradient_kept(const RenderArgs& A, const TriSample& T, const AxisTerms& X = AxisTerms{}) {
    float N[3];
    if constexpr (TAB) {
        GradTaps G;
        grad_taps_fetch(A.v, X, T.ix, T.iy, T.iz, G);
        grad_diff_finish(T, G, N);
    } else {
        grad_diff_kept(A.v, T, N);
    }
    if (A.invert_z) N[2] = -N[2];
    float r = rsqrtf(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    return make_float3(N[0] * r, N[1] * r, N[2] * r);
}

__device__ inline float3 gradient(const RenderArgs& A, const float p[3]) {
    TriSample T;
    (void)sample_keep(A.v, p, T);
    return gradient_kept(A, T);
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
    const uint32_t t = mk_xcd_row_tile(blockIdx.x, tiles_x, (A.height + TW - 1) / TW);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int WPR = TW / 8;  // waves per tile row
    x = (int)(t % tiles_x) * TW + (wave % WPR) * 8 + (lane & 7);
    y = (int)(t / tiles_x) * TW + (wave / WPR) * 8 + (lane >> 3);
    return x < A.width && y < A.height;
}

// Shared-memory carve-up common to the kernels. Per-label values a ray
// indexes by its sample's label live here rather than in the kernel
// arguments, whose per-lane indexing would be a memory round trip each.
constexpr int kLabelRows = 68;   // MK_NUM_LABELS rounded up
struct Smem {
    float4* aw;     // kLabelRows x (alpha, ww, wl, -)
    int4* lab;      // kLabelRows x (LUT offset, LUT length, accumulator slot, -)
    float4* lut;    // nlut_lds entries
    float* aobj;    // nslot x blockDim
    AxisTerms terms;  // per-axis address terms (kernels that stage them)
};

// LDS bytes of the carve-up ahead of the LUT entries.
constexpr size_t kLabelLdsBytes = kLabelRows * (sizeof(float4) + sizeof(int4));

__device__ inline void stage_labels(const RenderArgs& A, int4* lab) {
    for (int i = threadIdx.x; i < MK_NUM_LABELS; i += blockDim.x)
        lab[i] = make_int4(A.tab.off[i], A.tab.len[i], (int)A.slot_of[i], 0);
}

// Stages the label rows, the LUTs (stage_lut) and, after aobj_floats floats
// of per-label accumulators, the axis term tables (stage_terms).
__device__ inline Smem stage_shared(const RenderArgs& A, const AwTable& T, bool stage_lut, int aobj_floats = 0,
                                    bool stage_terms = false) {
    extern __shared__ float4 smem[];
    Smem s;
    s.aw = smem;
    s.lab = (int4*)(smem + kLabelRows);
    s.lut = smem + 2 * kLabelRows;
    s.aobj = (float*)(s.lut + A.nlut_lds);
    for (int i = threadIdx.x; i < MK_NUM_LABELS; i += blockDim.x) s.aw[i] = make_float4(T.alpha[i], T.ww[i], T.wl[i], 0.f);
    stage_labels(A, s.lab);
    if (stage_lut)
        for (int i = threadIdx.x; i < A.nlut_lds; i += blockDim.x) s.lut[i] = A.lut[i];
    s.terms = stage_terms ? stage_axis_terms(A.v, (uint32_t*)(s.aobj + aobj_floats)) : AxisTerms{};
    __syncthreads();
    return s;
}

__device__ inline float start_acc(const RenderArgs& A, float a_in) {
    int k0 = (int)(a_in / A.l1) - 2;
    k0 = k0 < 0 ? 0 : (k0 >= A.nsteps ? A.nsteps - 1 : k0);
    return A.steps[k0];
}

// skip_lattice: advances acc over the points acc + L1, acc + L1 + L1, ...
// that lie below a_skip (all inside one cell that cannot change the result),
// counting trips against the same budget as the marching loop; the same
// floats as adding one L1 at a time (lattice_skip, mk_lattice.h).
__device__ inline void skip_lattice(const RenderArgs& A, float& acc, float a_skip, int& it) {
    lattice_skip(acc, A.l1, A.rl1, a_skip, it, A.max_iter);
}

// Alpha of a sample taken at a refined step, as getNextStep corrects it
// (core/kernel.cu:84-125): 1 - (1 - a)^(step / L1).
__device__ inline float vr_step_alpha(float a, float step, float L1, float L4) {
    if (a > 0.f && step != L1) a = 1.0f - __builtin_amdgcn_exp2f((step == L4 ? 0.25f : 0.125f) * __builtin_amdgcn_logf(1.0f - a));
    return a;
}

// getNextStep's decision (core/kernel.cu:84-125) for a sample of corrected
// alpha a at acc: true when it backs up (acc and step changed, the sample is
// dropped); otherwise the sample is accepted, aPre becomes a and acc advances
// by the (possibly changed) step.
// Written as selects: the step is one of L1, L4, L8 exactly, and each branch's
// arithmetic is the same as in the nested form (acc - (L1 - L4) or
// acc - (L4 - L8) on backing up, acc + the new step otherwise).
__device__ inline bool vr_step_back(float& acc, float& step, float& aPre, float a, float L1, float L4, float L8) {
    const bool act = acc > 0.f;
    const bool hi = fmaxf(a, aPre) > 0.001f;
    const bool back = act && hi && step != L8;
    const float fine = step == L1 ? L4 : L8;      // the step after backing up from L1 or L4
    const float coarse = step == L8 ? L4 : L1;    // the step after a low sample
    const float next = (act && !hi) ? coarse : step;
    const float acc_back = acc - (step - fine);
    const float acc_on = acc + next;
    acc = back ? acc_back : acc_on;
    aPre = back ? aPre : a;
    step = back ? fine : next;
    return back;
}

// tracing's shading factor (core/kernel.cu:49-81) for normal N.
__device__ inline float vr_shade(const RenderArgs& A, const float3 N) {
    const float3 L = light_dir(A);
    const float d = N.x * L.x + N.y * L.y + N.z * L.z;
    float k = 0.35f;
    if (d > 0.f) {
        float d2 = d * d, d4 = d2 * d2;
        k += d * 0.8f + 0.16f * (d4 * d4);
    }
    return k;
}

// Front-to-back compositing of an accepted sample (core/kernel.cu:225-240).
__device__ inline void vr_composite(float4& sum, float& aAcc, float& slot, float a, float cr, float cg, float cb,
                                    float k) {
    const float w = (1.0f - aAcc) * a;
    sum.x += w * (cr * k);
    sum.y += w * (cg * k);
    sum.z += w * (cb * k);
    sum.w += w * (a * k);
    slot += w;
    aAcc += w;
}

// vr_composite with the products colour * k already formed (the same
// roundings: each product is one multiply in either form).
__device__ inline void vr_composite_k(float4& sum, float& aAcc, float& slot, float a, float ckr, float ckg, float ckb,
                                      float cka) {
    const float w = (1.0f - aAcc) * a;
    sum.x += w * ckr;
    sum.y += w * ckg;
    sum.z += w * ckb;
    sum.w += w * cka;
    slot += w;
    aAcc += w;
}

// d_render (core/kernel.cu:133-264) for pixel (x, y); tid indexes the
// per-label accumulators of a block of BS threads. TOUCH: measurement variant
// that skips nothing and marks every voxel the reference rule reads.
template <int BS, bool MASK, bool LDS_LUT, bool TOUCH, bool TAB = false>
__device__ inline void vr_pixel(const RenderArgs& A, const Smem& s, int x, int y, int tid) {

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
        int trips = 0, it_queued = -1;
        // finished first-pass waves as read at the previous poll: a poll's
        // load is used 16 trips later, so its round trip overlaps the ray's
        // own loads instead of stalling the wave
        uint32_t seen = 0;
        for (int it = 0; it < A.max_iter && acc <= kAccEnd && acc <= a_out; ++it) {
            // a ray still marching after the budget, or when most of the
            // first pass's waves have finished (this wave is in its tail), is
            // queued for the second pass; all lanes of a wave take either
            // exit on the same trip
            if (!TOUCH && A.cont.cap > 0) {
                if (trips == A.cont.budget) { it_queued = it; break; }
                if ((trips & 15) == 15) {
                    if (seen >= A.cont.tail_waves) { it_queued = it; break; }
                    seen = __hip_atomic_load(&A.cont.counters[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                ++trips;
            }
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
                    const uint32_t e = A.bskip[b];
                    cur_macro_empty = !(e & MK_BSKIP_MACRO);
                    cur_occ = (e & MK_BSKIP_OCC) != 0;
                    cur_label = MASK ? (int)(e & MK_BSKIP_LABEL) : 0;
                }
                empty_shift = cur_macro_empty ? 5 : (cur_occ ? 0 : 3);
            }
            // The empty and the sampled cases each run getNextStep
            // (kernel.cu:84-125) on their own: the sample's state (corners,
            // colour, label) then never has to meet the empty case's zeros.
            if (!empty_shift) {
                // the corner loads are issued before the mask label is read,
                // so that both are in flight together
                TriFetch fr;
                if constexpr (TAB) sample_fetch(A.v, s.terms, p, fr);
                else sample_fetch(A.v, p, fr);
                int label = 0;
                if (MASK) label = (TOUCH || cur_label == 0xFF) ? A.v.mask[mk_mask_offset(A.v, cx, cy, cz)] : cur_label;
                if (TOUCH) {
                    if (MASK) touch_mask(A, cx, cy, cz);
                    touch_footprint(A, p[0], p[1], p[2]);
                }
                const float4 aw = s.aw[label];
                const int4 lb = s.lab[label];
                TriSample tri;
                float t = sample_finish(fr, tri);
                t = (t - aw.z) / aw.y + 0.5f;
                t = t > 1.0f ? 1.0f : t;
                const float4 col = lut_sample<LDS_LUT>(A, s.lut, lb.x, lb.y, t);
                const float a = vr_step_alpha(col.w, step, L1, L4);
                if (vr_step_back(acc, step, aPre, a, L1, L4, L8)) continue;
                if (a > 0.0005f) {
                    const float cap = aw.x;
                    float* slot = MASK ? &s.aobj[lb.z * BS + tid] : &aobj0;
                    if (*slot < cap) {
                        // tracing (kernel.cu:49-81)
                        if (TOUCH) touch_gradient(A, p);
                        const float k = vr_shade(A, gradient_kept<TAB>(A, tri, s.terms));
                        vr_composite(sum, aAcc, *slot, a, col.x, col.y, col.z, k);
                    }
                }
            } else {
                // alpha 0: no correction, no composite
                if (vr_step_back(acc, step, aPre, 0.f, L1, L4, L8)) continue;
            }
            if (aAcc > A.exit_alpha) break;
            // steady state in an empty cell: every further lattice point inside
            // it is accepted with alpha 0 and changes nothing but acc
            if (empty_shift && step == L1) skip_lattice(A, acc, cell_exit(A, g, cx, cy, cz, empty_shift), it);
        }
        if (!TOUCH && it_queued >= 0) {
            // one queue slot per lane, one atomic per wave and end; the queue
            // holds every pixel of the frame, so it cannot overflow. Rays with
            // a long way left fill it from the front, the others from the
            // back, and the second pass takes the front first: the longest
            // work starts first.
            const VrCont& C = A.cont;
            const bool lng = (a_out - acc) >= C.long_steps * L1;
            const uint64_t m = __ballot(lng);
            const uint64_t all = __ballot(1);
            const uint64_t mine = lng ? m : (all & ~m);
            const int lane = __lane_id();
            const int leader = __ffsll((long long)mine) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&C.counters[lng ? 0 : 2], (uint32_t)__popcll(mine));
            base = (uint32_t)__shfl((int)base, leader);
            const uint32_t k = base + (uint32_t)__popcll(mine & ((1ull << lane) - 1ull));
            const uint32_t r = lng ? k : (uint32_t)C.cap - 1u - k;
            C.pix[r] = y * A.width + x;
            C.acc[r] = acc;
            C.step[r] = step;
            C.apre[r] = aPre;
            C.aacc[r] = aAcc;
            C.it[r] = it_queued;
            C.sum[r] = sum;
            if (MASK) {
                for (int k = 0; k < A.nslot; ++k) C.aobj[(size_t)k * C.cap + r] = s.aobj[k * BS + tid];
            } else {
                C.aobj[r] = aobj0;
            }
            return;
        }
    }
    if (sum.x == 0.f && sum.y == 0.f && sum.z == 0.f && sum.w == 0.f) sum = bg4(A);
    store_rgb(A, x, y, sum);
}

// Second pass over the queued rays: G lanes per ray. While the step stays the
// same, the positions of a ray's next samples are known (the reference's float
// accumulation acc + step + step ...), so lane j fetches and evaluates the
// j-th of them (and its shading, if it may composite). Within such a run a
// sample depends on the ones before it only through aPre (the alpha of the
// last sample in the VOI before it), the step and the early exit; so each lane
// finds its own aPre and the event its sample causes in the reference's loop
// (core/kernel.cu:133-264: end of the ray, backing up, or a step change after
// it), the first event ends the run, and the samples that may composite are
// composited in order up to there with the same per-sample code as vr_pixel.
// The frame is the same, bit for bit, as one pass's. gobj: this group's
// per-label accumulators in LDS (MASK).
template <bool MASK, bool LDS_LUT, int G, bool TAB = false>
__device__ inline void vr_tail(const RenderArgs& A, const Smem& s, float* gobj) {
    static_assert(G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "group size");
    const int lane = __lane_id();
    const int j = lane & (G - 1), base = lane - j;
    const uint64_t gbits = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const float L1 = A.l1, L4 = L1 / 4.0f, L8 = L1 / 8.0f;
    const VrCont& C = A.cont;
    const uint32_t n_front = C.counters[0], total = n_front + C.counters[2];
    bool have = false, fin = false;
    int pix = 0;
    float u = 0.f, v = 0.f, a_out = 0.f;
    RayGeom g;
    float acc = 0.f, step = L1, aPre = 0.f, aAcc = 0.f, aobj0 = 0.f;
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
    int it = 0;
    // this lane's brick cache
    uint32_t cur_brick = 0xFFFFFFFFu;
    int cur_macro_empty = 0, cur_occ = 1, cur_label = 0xFF;
    for (;;) {
        if (!have && !fin) {
            uint32_t r = 0;
            if (j == 0) r = atomicAdd(&C.counters[1], 1u);
            r = (uint32_t)__shfl((int)r, base);
            if (r >= total || total > (uint32_t)C.cap) {
                fin = true;
            } else {
                if (r >= n_front) r = (uint32_t)C.cap - 1u - (r - n_front);
                pix = C.pix[r];
                const int x = pix % A.width, y = pix / A.width;
                u = ((float)x - (float)A.width / 2.0f - A.xT) / (float)A.height;
                v = ((float)y - (float)A.height / 2.0f - A.yT) / (float)A.height;
                float a_in;
                (void)ray_interval(A, u, v, g, a_in, a_out);
                acc = C.acc[r];
                step = C.step[r];
                aPre = C.apre[r];
                aAcc = C.aacc[r];
                it = C.it[r];
                sum = C.sum[r];
                if (MASK) {
                    for (int k = 0; k < A.nslot; ++k) gobj[k] = C.aobj[(size_t)k * C.cap + r];
                } else {
                    aobj0 = C.aobj[r];
                }
                have = true;
            }
        }
        if (__ballot(have) == 0ull) break;
        if (!have) continue;

        // this lane's sample: the j-th position from acc at the current step
        float accj = acc;
        for (int q = 0; q < j; ++q) accj += step;
        float p[3];
        ray_pos(A, u, v, accj, p);
        int ix, iy, iz;
        const bool inv = in_voi(A, p, ix, iy, iz);
        // for a sample that may composite: the colour-times-shade products,
        // the label's cap and its accumulator slot. A sample in an empty cell
        // reads alpha 0; the lattice points vr_pixel would skip after it are
        // visited one by one, with the same result.
        bool cand = false;
        int slot_i = 0;
        float a = 0.f, kr = 0.f, kg = 0.f, kb = 0.f, ka = 0.f, cap = 0.f;
        if (inv) {
            const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1), cz = mk_clampi(iz, 0, A.v.nz - 1);
            const uint32_t b = mk_brick_index(A.v, cx, cy, cz);
            if (b != cur_brick) {
                cur_brick = b;
                const uint32_t e = A.bskip[b];
                cur_macro_empty = !(e & MK_BSKIP_MACRO);
                cur_occ = (e & MK_BSKIP_OCC) != 0;
                cur_label = MASK ? (int)(e & MK_BSKIP_LABEL) : 0;
            }
            if (!(cur_macro_empty || !cur_occ)) {
                // corners first, then the mask label: both in flight together
                TriFetch fr;
                if constexpr (TAB) sample_fetch(A.v, s.terms, p, fr);
                else sample_fetch(A.v, p, fr);
                int label = 0;
                if (MASK) label = cur_label == 0xFF ? A.v.mask[mk_mask_offset(A.v, cx, cy, cz)] : cur_label;
                const float4 aw = s.aw[label];
                const int4 lb = s.lab[label];
                TriSample tri;
                float t = sample_finish(fr, tri);
                t = (t - aw.z) / aw.y + 0.5f;
                t = t > 1.0f ? 1.0f : t;
                const float4 col = lut_sample<LDS_LUT>(A, s.lut, lb.x, lb.y, t);
                a = vr_step_alpha(col.w, step, L1, L4);
                // the label's accumulator only grows: one at its cap now
                // stays there, and its samples never composite
                const float held = MASK ? gobj[lb.z] : aobj0;
                if (a > 0.0005f && held < aw.x) {
                    const float k = vr_shade(A, gradient_kept<TAB>(A, tri, s.terms));
                    kr = col.x * k;
                    kg = col.y * k;
                    kb = col.z * k;
                    ka = a * k;
                    cap = aw.x;
                    slot_i = MASK ? lb.z : 0;
                    cand = true;
                }
            }
        }

        // aPre at this sample: the alpha of the last sample in the VOI before
        // it in this run, else the ray's
        const uint64_t vb = (__ballot(inv) >> base) & gbits;
        const uint64_t below = vb & ((1ull << j) - 1ull);
        const int src = below ? base + 63 - __clzll((long long)below) : lane;
        const float a_prev = __shfl(a, src);
        const float apre_j = below ? a_prev : aPre;
        // this sample's event in the reference's loop: 0 none, 1 the loop
        // ends before it, 2 getNextStep backs up at it, 3 it is accepted and
        // then the step changes
        int ev = 0;
        if (!(it + j < A.max_iter && accj <= kAccEnd && accj <= a_out)) {
            ev = 1;
        } else if (inv && accj > 0.f) {
            const bool hi = fmaxf(a, apre_j) > 0.001f;
            if (step == L8) ev = hi ? 0 : 3;
            else if (step == L4) ev = hi ? 2 : 3;
            else ev = hi ? 2 : 0;
        }
        const uint64_t eb = (__ballot(ev != 0) >> base) & gbits;
        const int e = eb ? __ffsll((long long)eb) - 1 : G;
        const int ev_e = __shfl(ev, base + (e < G ? e : G - 1));
        const int kind = e < G ? ev_e : 0;

        // composite, in order, the samples that may do so up to the event
        const int lim = kind == 3 ? e + 1 : e;
        uint64_t cb = (__ballot(cand) >> base) & gbits;
        if (lim < 64) cb &= (1ull << lim) - 1ull;
        bool done = kind == 1;
        while (cb) {
            const int q = __ffsll((long long)cb) - 1;
            cb &= cb - 1ull;
            const float c_q = __shfl(cap, base + q);
            const float r_q = __shfl(kr, base + q);
            const float g_q = __shfl(kg, base + q);
            const float b_q = __shfl(kb, base + q);
            const float w_q = __shfl(ka, base + q);
            const float a_q = __shfl(a, base + q);
            const int s_q = __shfl(slot_i, base + q);
            float& slot = MASK ? gobj[s_q] : aobj0;
            if (slot < c_q) vr_composite_k(sum, aAcc, slot, a_q, r_q, g_q, b_q, w_q);
            if (aAcc > A.exit_alpha) {
                done = true;
                break;
            }
        }

        if (!done) {
            // the state after the run, as the reference's loop leaves it
            const int last = kind == 0 ? G - 1 : e;
            const float acc_l = __shfl(accj, base + last);
            const float a_l = __shfl(a, base + last);
            const float apre_l = __shfl(apre_j, base + last);
            const uint64_t vbl = vb & (last >= 63 ? ~0ull : ((1ull << (last + 1)) - 1ull));
            const int srcl = vbl ? base + 63 - __clzll((long long)vbl) : lane;
            const float a_vl = __shfl(a, srcl);
            if (kind == 0) {
                // every sample accepted (out-of-VOI ones only advance acc)
                acc = acc_l + step;
                if (vbl) aPre = a_vl;
                it += G;
            } else if (kind == 2) {
                // backing up at sample e, which is dropped
                if (step == L1) {
                    acc = acc_l - (L1 - L4);
                    step = L4;
                } else {
                    acc = acc_l - (L4 - L8);
                    step = L8;
                }
                aPre = apre_l;
                it += e + 1;
            } else {
                // sample e accepted, then the step changes
                step = (step == L8) ? L4 : L1;
                aPre = a_l;
                acc = acc_l + step;
                it += e + 1;
            }
        }
        if (done) {
            if (j == 0) {
                float4 o = sum;
                if (o.x == 0.f && o.y == 0.f && o.z == 0.f && o.w == 0.f) o = bg4(A);
                store_rgb(A, pix % A.width, pix / A.width, o);
            }
            have = false;
        }
    }
}

// MIP's march over one ray, one lattice point at a time (d_renderMIP,
// core/kernel.cu:266-364): a cell whose bound cannot exceed the running
// maximum leaves it unchanged and is skipped.
template <bool MASK, bool TOUCH, bool TAB = false>
__device__ inline void mip_march_single(const RenderArgs& A, const Smem& s, float u, float v, const RayGeom& g,
                                        float acc, float a_out, float& aAcc) {
    const float L1 = A.l1;
    uint32_t cur_brick = 0xFFFFFFFFu;
    float cur_bm = INFINITY, cur_bb = INFINITY;
    int cur_label = 0xFF;
    for (int it = 0; it < A.max_iter && acc <= kAccEnd && acc <= a_out; ++it) {
        float p[3];
        ray_pos(A, u, v, acc, p);
        int ix, iy, iz;
        if (in_voi(A, p, ix, iy, iz)) {
            const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1), cz = mk_clampi(iz, 0, A.v.nz - 1);
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
                int label = MASK ? ((TOUCH || cur_label == 0xFF) ? A.v.mask[mk_mask_offset(A.v, cx, cy, cz)] : cur_label) : 0;
                if (TOUCH && MASK) touch_mask(A, cx, cy, cz);
                const float4 aw = s.aw[label];
                if (aw.x > 0.f) {
                    if (TOUCH) touch_footprint(A, p[0], p[1], p[2]);
                    float t;
                    if constexpr (TAB) {
                        TriSample tri;
                        t = sample_keep<true>(A.v, s.terms, p, tri);
                    } else {
                        t = mk_sample_hu(A.v, p[0], p[1], p[2]);
                    }
                    t = (t - aw.z) / aw.y + 0.5f;
                    if (aAcc < t) aAcc = t;
                }
            }
            acc += L1;
            if (skip_shift) skip_lattice(A, acc, cell_exit(A, g, cx, cy, cz, skip_shift), it);
        } else {
            acc += L1;
        }
    }
}

// The same march kMipGroup lattice points at a time. The result is the
// maximum of the samples, which does not depend on the order they are taken
// in: the group's bricks are looked up first, then all its fetches are issued,
// then they are folded into the maximum. Deciding a skip against the maximum
// from before the group only skips less, so the result is the same as one
// point at a time.
template <bool MASK, bool TOUCH, bool TAB = false>
__device__ inline void mip_march_grouped(const RenderArgs& A, const Smem& s, float u, float v, const RayGeom& g,
                                         float acc, float a_out, float& aAcc) {
    constexpr int kMipGroup = 4;
    const float L1 = A.l1;
    uint32_t cur_brick = 0xFFFFFFFFu;
    float cur_bm = INFINITY, cur_bb = INFINITY;
    int cur_label = 0xFF;
    int it = 0;
    while (it < A.max_iter && acc <= kAccEnd && acc <= a_out) {
        // 1: the group's points, their bricks' skip entries and labels
        float pp[kMipGroup][3];
        int lab[kMipGroup];
        bool want[kMipGroup];
        int n = 0, last_shift = 0, lx = 0, ly = 0, lz = 0;
        float acc_j = acc;
#pragma unroll
        for (int j = 0; j < kMipGroup; ++j) {
            want[j] = false;
            lab[j] = 0;
            // the loop's bound, point by point: a contiguous run
            if (n == j && it + j < A.max_iter && acc_j <= kAccEnd && acc_j <= a_out) {
                n = j + 1;
                last_shift = 0;
                ray_pos(A, u, v, acc_j, pp[j]);
                int ix, iy, iz;
                if (in_voi(A, pp[j], ix, iy, iz)) {
                    const int cx = mk_clampi(ix, 0, A.v.nx - 1), cy = mk_clampi(iy, 0, A.v.ny - 1),
                              cz = mk_clampi(iz, 0, A.v.nz - 1);
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
                        lab[j] = MASK ? ((TOUCH || cur_label == 0xFF) ? A.v.mask[mk_mask_offset(A.v, cx, cy, cz)] : cur_label) : 0;
                        if (TOUCH && MASK) touch_mask(A, cx, cy, cz);
                        want[j] = true;
                    }
                    last_shift = skip_shift;
                    lx = cx;
                    ly = cy;
                    lz = cz;
                }
                acc_j += L1;
            }
        }
        // 2: every wanted sample's corners, all in flight together
        TriFetch fr[kMipGroup];
#pragma unroll
        for (int j = 0; j < kMipGroup; ++j) {
            if (want[j] && s.aw[lab[j]].x > 0.f) {
                if (TOUCH) touch_footprint(A, pp[j][0], pp[j][1], pp[j][2]);
                if constexpr (TAB) sample_fetch(A.v, s.terms, pp[j], fr[j]);
                else sample_fetch(A.v, pp[j], fr[j]);
            } else {
                want[j] = false;
            }
        }
        // 3: the maximum
#pragma unroll
        for (int j = 0; j < kMipGroup; ++j) {
            if (want[j]) {
                const float4 aw = s.aw[lab[j]];
                TriSample tri;
                float t = sample_finish(fr[j], tri);
                t = (t - aw.z) / aw.y + 0.5f;
                if (aAcc < t) aAcc = t;
            }
        }
        acc = acc_j;
        it += n;
        // the last point's cell, skipped as one point at a time would
        if (last_shift) skip_lattice(A, acc, cell_exit(A, g, lx, ly, lz, last_shift), it);
    }
}

// d_renderMIP (core/kernel.cu:266-364) for pixel (x, y). With a mask the
// march takes four lattice points at a time (measured faster); without one a
// group's fetches did not stay in flight together, and one point at a time is
// faster.
template <bool MASK, bool TOUCH, bool TAB = false>
__device__ inline void mip_pixel(const RenderArgs& A, const Smem& s, int x, int y) {
    const float u = ((float)x - (float)A.width / 2.0f - A.xT) / (float)A.height;
    const float v = ((float)y - (float)A.height / 2.0f - A.yT) / (float)A.height;
    float aAcc = 0.f;
    float a_in, a_out;
    RayGeom g;
    if (ray_interval(A, u, v, g, a_in, a_out)) {
        if constexpr (MASK) mip_march_grouped<MASK, TOUCH, TAB>(A, s, u, v, g, start_acc(A, a_in), a_out, aAcc);
        else mip_march_single<MASK, TOUCH, TAB>(A, s, u, v, g, start_acc(A, a_in), a_out, aAcc);
    }
    float4 c = aAcc <= 0.f ? bg4(A) : make_float4(aAcc, aAcc, aAcc, aAcc);
    store_rgb(A, x, y, c);
}

// d_renderSurface (core/kernel.cu:366-469) for pixel (x, y).
template <bool TOUCH>
__device__ inline void surface_pixel(const RenderArgs& A, int x, int y) {
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
        for (int it = 0; it < A.max_iter && acc <= kAccEnd && acc <= a_out; ++it) {
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
                if (skip_shift) skip_lattice(A, acc, cell_exit(A, g, cx, cy, cz, skip_shift), it);
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


}  // namespace mkr
