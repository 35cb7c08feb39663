// mk_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the MonkeyGL render path (reference core/kernel.cu and
// the texture semantics of core/CuDataManager.cu). It is the checker used by
// tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, and
// is never linked into, imported by or called from the product library.
//
// Every routine is a template on the arithmetic type T:
//   T = double : the FP64 oracle (exact trilinear weights);
//   T = float  : the CPU baseline, in the reference's float op order; with
//                weights8 = 1 it also quantises filter weights to 8
//                fractional bits as CUDA texture units do (public CUDA
//                Programming Guide, texture fetching appendix).
// Built with -ffp-contract=off so the float instance is plain IEEE arithmetic.
//
// Pinning (what fixes each function independently of this file), see tests:
//   tex3_linear / tex1_linear: ramp volumes give analytic values; voxel
//     centres give raw voxels; clamp at edges.
//   vr_pixel: constant volume => ambient-only shading closed form; single
//     opaque slab => analytic compositing; empty TF => background.
//   mip_pixel: ramp volume => analytic maximum.
//   surface_pixel: step edge => first-hit gray closed form.
//   plane_pixel: axis-aligned planes reduce to direct voxel reads; slab
//     Average/MIP/MinIP of a z-ramp are closed forms.
//   cpr_pixel: a straight centreline reduces to an MPR row.
//   mc_*: all 256 single-cube cases, watertightness, Euler characteristic,
//     vertices on edges at the iso value.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../include/monkeygl_hip.h"
#include "../monkeygl_amd/csrc/mk_mc_table.h"

namespace {

struct Volume {
    const int16_t* v;
    const uint8_t* mask;  // may be null
    int nx, ny, nz;
    uint8_t* touched;     // optional per-voxel footprint map (V_touched)
    inline int16_t at(int x, int y, int z) const {
        return v[((size_t)z * ny + y) * nx + x];
    }
    inline void touch(int x, int y, int z) const {
        if (touched) touched[((size_t)z * ny + y) * nx + x] = 1;
    }
};

struct Luts {
    const float* const* lut;  // MK_NUM_LABELS pointers (null = all-zero LUT)
    const int* len;
};

template <typename T> static inline T clampi_t(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// cudaReadModeNormalizedFloat for signed 16-bit: v / 32767, clamped at -1
// (core/CuDataManager.cu:71).
template <typename T> static inline T norm16(int16_t v) {
    T r = T(v) / T(32767);
    return r < T(-1) ? T(-1) : r;
}

template <typename T> static inline T quant8(T a) {
    return std::nearbyint(a * T(256)) / T(256);
}

// Texel split of one normalized coordinate: x_t = f*N - 0.5, i = floor(x_t),
// a = x_t - i (core/CuDataManager.cu:60-68, normalized coords + linear filter).
template <typename T, bool Q8> static inline void split(T f, int n, int& i0, int& i1, T& a) {
    T xt = f * T(n) - T(0.5);
    T fl = std::floor(xt);
    a = xt - fl;
    if (Q8) a = quant8(a);
    // clamp before the int conversion so huge coordinates stay defined
    fl = clampi_t<T>(fl, T(-2), T(n + 1));
    int i = (int)fl;
    i0 = clampi(i, 0, n - 1);
    i1 = clampi(i + 1, 0, n - 1);
}

// tex3D<float> of the volume texture: normalized coordinates, linear filter,
// clamp addressing, normalized-float read (core/CuDataManager.cu:38-72).
template <typename T, bool Q8> static T tex3_linear(const Volume& vol, T fx, T fy, T fz) {
    int x0, x1, y0, y1, z0, z1;
    T ax, ay, az;
    split<T, Q8>(fx, vol.nx, x0, x1, ax);
    split<T, Q8>(fy, vol.ny, y0, y1, ay);
    split<T, Q8>(fz, vol.nz, z0, z1, az);
    if (vol.touched) {
        vol.touch(x0, y0, z0); vol.touch(x1, y0, z0); vol.touch(x0, y1, z0); vol.touch(x1, y1, z0);
        vol.touch(x0, y0, z1); vol.touch(x1, y0, z1); vol.touch(x0, y1, z1); vol.touch(x1, y1, z1);
    }
    T c000 = norm16<T>(vol.at(x0, y0, z0)), c100 = norm16<T>(vol.at(x1, y0, z0));
    T c010 = norm16<T>(vol.at(x0, y1, z0)), c110 = norm16<T>(vol.at(x1, y1, z0));
    T c001 = norm16<T>(vol.at(x0, y0, z1)), c101 = norm16<T>(vol.at(x1, y0, z1));
    T c011 = norm16<T>(vol.at(x0, y1, z1)), c111 = norm16<T>(vol.at(x1, y1, z1));
    T c00 = (T(1) - ax) * c000 + ax * c100;
    T c10 = (T(1) - ax) * c010 + ax * c110;
    T c01 = (T(1) - ax) * c001 + ax * c101;
    T c11 = (T(1) - ax) * c011 + ax * c111;
    T c0 = (T(1) - ay) * c00 + ay * c10;
    T c1 = (T(1) - ay) * c01 + ay * c11;
    return (T(1) - az) * c0 + az * c1;
}

// Point-sampled read for nearest MPR/CPR (new mode): voxel floor(f*N),
// clamped, normalized-float read.
template <typename T> static T tex3_nearest(const Volume& vol, T fx, T fy, T fz) {
    T tx = std::floor(fx * T(vol.nx)), ty = std::floor(fy * T(vol.ny)), tz = std::floor(fz * T(vol.nz));
    int x = clampi((int)clampi_t<T>(tx, T(-1), T(vol.nx)), 0, vol.nx - 1);
    int y = clampi((int)clampi_t<T>(ty, T(-1), T(vol.ny)), 0, vol.ny - 1);
    int z = clampi((int)clampi_t<T>(tz, T(-1), T(vol.nz)), 0, vol.nz - 1);
    vol.touch(x, y, z);
    return norm16<T>(vol.at(x, y, z));
}

// tex1D<float4> of a transfer-function LUT: normalized coordinates, linear
// filter, clamp (core/CuDataManager.cu:136-145). A missing LUT reads zero.
template <typename T, bool Q8> static void tex1_linear(const Luts& L, int label, T s, T out[4]) {
    const float* lut = L.lut[label];
    int n = L.len[label];
    if (!lut || n <= 0) { out[0] = out[1] = out[2] = out[3] = T(0); return; }
    if (!(s >= T(-1))) s = T(-1);  // NaN or -inf read entry 0, as clamp does
    int i0, i1;
    T a;
    split<T, Q8>(s, n, i0, i1, a);
    for (int c = 0; c < 4; ++c) out[c] = (T(1) - a) * T(lut[i0 * 4 + c]) + a * T(lut[i1 * 4 + c]);
}

// helper_math-style saturate; NaN maps to 0.
template <typename T> static inline T sat(T x) { return x > T(0) ? (x < T(1) ? x : T(1)) : T(0); }

// rgbaFloatToInt (core/kernel.cu:40-47) written as three bytes.
template <typename T> static inline void write_rgb(uint8_t* o, const T c[4]) {
    o[0] = (uint8_t)(unsigned)(sat(c[0]) * T(255));
    o[1] = (uint8_t)(unsigned)(sat(c[1]) * T(255));
    o[2] = (uint8_t)(unsigned)(sat(c[2]) * T(255));
}

template <typename T> struct Ray {
    float u, v;
    float M[9];
    float mls[3];
    int invz;
    // core/kernel.cu:197-210 position of the sample at accumulated length acc.
    // Positions are float in every instance, as the reference computes them:
    // the reference defines its sample lattice by float accumulation
    // (kernel.cu:169-172, 213, 238) and its sample positions by float
    // arithmetic, and in views along an axis those positions fall exactly on
    // voxel faces, where only their float value decides the VOI test and the
    // mask label of a whole column of samples. The multiply-adds are fused,
    // the dot product left to right, as CUDA's default contraction compiles
    // the reference's expressions (and as the HIP kernels write them). The
    // FP64 oracle evaluates everything after the position in double.
    inline void pos(float acc, double scale, T p[3]) const {
        const float fy = (float)(((double)acc - 0.866) * scale);
        const float x = std::fma(v, M[2], std::fma(fy, M[1], u * M[0]));
        const float y = std::fma(v, M[5], std::fma(fy, M[4], u * M[3]));
        const float z = std::fma(v, M[8], std::fma(fy, M[7], u * M[6]));
        float pz = std::fma(z, mls[2], 0.5f);
        if (invz) pz = 1.0f - pz;
        p[0] = (T)std::fma(x, mls[0], 0.5f);
        p[1] = (T)std::fma(y, mls[1], 0.5f);
        p[2] = (T)pz;
    }
};

template <typename T> static inline bool in_voi(const mk_voi& voi, const Volume& vol, const T p[3], int& ix, int& iy, int& iz) {
    // core/kernel.cu:208-211: truncation toward zero of pos*dim, a float
    // product there (pos is float3, the extent converts to float); p holds
    // the float position (Ray::pos), so every instance takes the reference's
    // voxel index and with it its VOI decision and mask label
    const float fx = (float)p[0] * (float)vol.nx, fy = (float)p[1] * (float)vol.ny, fz = (float)p[2] * (float)vol.nz;
    // guard the int conversion; anything this far out fails the test anyway
    if (!(fx > -1e9f && fx < 1e9f && fy > -1e9f && fy < 1e9f && fz > -1e9f && fz < 1e9f)) return false;
    ix = (int)fx; iy = (int)fy; iz = (int)fz;
    return !(ix < voi.left || ix > voi.right || iy < voi.posterior || iy > voi.anterior || iz < voi.head || iz > voi.foot);
}

template <typename T> static inline uint8_t mask_at(const Volume& vol, int ix, int iy, int iz) {
    if (!vol.mask) return 0;
    int x = clampi(ix, 0, vol.nx - 1), y = clampi(iy, 0, vol.ny - 1), z = clampi(iz, 0, vol.nz - 1);
    uint8_t l = vol.mask[((size_t)z * vol.ny + y) * vol.nx + x];
    return l > MK_MAX_OBJECTS ? (uint8_t)MK_MAX_OBJECTS : l;
}

template <typename T> static inline void normalize3(T n[3]) {
    T d = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    T r = T(1) / std::sqrt(d);  // d == 0 gives inf and then NaN, like rsqrtf
    n[0] *= r; n[1] *= r; n[2] *= r;
}

// Central-difference gradient at one voxel distance in texture space
// (core/kernel.cu:60-67).
template <typename T, bool Q8> static void gradient(const Volume& vol, const T p[3], const T sv[3], int invz, T N[3]) {
    N[0] = tex3_linear<T, Q8>(vol, p[0] + sv[0], p[1], p[2]) - tex3_linear<T, Q8>(vol, p[0] - sv[0], p[1], p[2]);
    N[1] = tex3_linear<T, Q8>(vol, p[0], p[1] + sv[1], p[2]) - tex3_linear<T, Q8>(vol, p[0], p[1] - sv[1], p[2]);
    N[2] = tex3_linear<T, Q8>(vol, p[0], p[1], p[2] + sv[2]) - tex3_linear<T, Q8>(vol, p[0], p[1], p[2] - sv[2]);
    if (invz) N[2] = -N[2];
    normalize3(N);
}

template <typename T> static Ray<T> make_ray(const mk_render_params& P, int x, int y) {
    Ray<T> r;
    // core/kernel.cu:161-162, in float as there
    r.u = ((float)x - (float)P.width / 2.0f - P.x_translate) / (float)P.height;
    r.v = ((float)y - (float)P.height / 2.0f - P.y_translate) / (float)P.height;
    for (int i = 0; i < 9; ++i) r.M[i] = P.transform[i];
    for (int i = 0; i < 3; ++i) r.mls[i] = P.max_len_spacing[i];
    r.invz = P.invert_z;
    return r;
}

template <typename T> static void light_dir(const mk_render_params& P, T L[3]) {
    // normalize(mul(M, (0,1,0))) (core/kernel.cu:166-167)
    L[0] = T(P.transform[1]); L[1] = T(P.transform[4]); L[2] = T(P.transform[7]);
    normalize3(L);
}

// d_render (core/kernel.cu:133-264) for one pixel, with tracing (49-81) and
// getNextStep (84-125).
template <typename T, bool Q8>
static void vr_pixel(const mk_render_params& P, const Volume& vol, const Luts& L, int x, int y, uint8_t* out) {
    Ray<T> ray = make_ray<T>(P, x, y);
    T dirL[3];
    light_dir<T>(P, dirL);
    T sv[3] = {T(P.spacing_voxel[0]), T(P.spacing_voxel[1]), T(P.spacing_voxel[2])};
    const double scale = (double)P.scale;

    const float L1 = 1.0f / (float)vol.nz * P.step_scale;
    const float L4 = L1 / 4.0f;
    const float L8 = L1 / 8.0f;
    const T exitA = T(P.exit_alpha);
    float step = L1;
    T aObj[MK_NUM_LABELS];
    for (int i = 0; i < MK_NUM_LABELS; ++i) aObj[i] = T(0);
    T aAcc = T(0), aPre = T(0);
    float acc = 0.0f;
    T sum[4] = {T(0), T(0), T(0), T(0)};

    while ((double)acc < 1.732) {
        T p[3];
        ray.pos(acc, scale, p);
        int ix, iy, iz;
        if (!in_voi<T>(P.voi, vol, p, ix, iy, iz)) { acc += step; continue; }
        uint8_t label = mask_at<T>(vol, ix, iy, iz);
        const mk_alpha_wwwl& aw = P.alpha_wwwl[label];
        T t = T(32768) * tex3_linear<T, Q8>(vol, p[0], p[1], p[2]);
        t = (T)((double)((t - T(aw.wl)) / T(aw.ww)) + 0.5);
        if (t > T(1)) t = T(1);
        T col[4];
        tex1_linear<T, Q8>(L, label, t, col);
        T a = col[3];
        // getNextStep
        if (step == L4) a = T(1) - std::pow(T(1) - a, T(0.25f));
        else if (step == L8) a = T(1) - std::pow(T(1) - a, T(0.125f));
        if (acc > 0.0f) {
            if (std::max(a, aPre) > T(0.001f)) {
                if (step == L1) { acc -= (L1 - L4); step = L4; continue; }
                else if (step == L4) { acc -= (L4 - L8); step = L8; continue; }
            } else {
                step = (step == L8) ? L4 : L1;
            }
        }
        aPre = a;
        acc += step;
        col[3] = a;
        if (a > T(0.0005f) && aObj[label] < T(aw.alpha)) {
            // tracing
            T N[3];
            gradient<T, Q8>(vol, p, sv, ray.invz, N);
            T d = N[0] * dirL[0] + N[1] * dirL[1] + N[2] * dirL[2];
            T cl[4];
            for (int c = 0; c < 4; ++c) cl[c] = col[c] * T(0.35f);
            if (d > T(0)) {
                T k = d * T(0.8f) + T(0.16f) * std::pow(d, T(8));
                for (int c = 0; c < 4; ++c) cl[c] += col[c] * k;
            }
            T w = (T(1) - aAcc) * a;
            for (int c = 0; c < 4; ++c) sum[c] = sum[c] + w * cl[c];
            aObj[label] += (T(1) - aAcc) * a;
            aAcc += (T(1) - aAcc) * a;
        }
        if (aAcc > exitA) break;
    }
    if (sum[0] == T(0) && sum[1] == T(0) && sum[2] == T(0) && sum[3] == T(0))
        for (int c = 0; c < 4; ++c) sum[c] = T(P.background[c]);
    write_rgb<T>(out, sum);
}

// d_renderMIP (core/kernel.cu:266-364).
template <typename T, bool Q8>
static void mip_pixel(const mk_render_params& P, const Volume& vol, int x, int y, uint8_t* out) {
    Ray<T> ray = make_ray<T>(P, x, y);
    const double scale = (double)P.scale;
    const float step = 1.0f / (float)vol.nz * P.step_scale;
    T aAcc = T(0);
    float acc = 0.0f;
    while ((double)acc < 1.732) {
        T p[3];
        ray.pos(acc, scale, p);
        int ix, iy, iz;
        if (!in_voi<T>(P.voi, vol, p, ix, iy, iz)) { acc += step; continue; }
        uint8_t label = mask_at<T>(vol, ix, iy, iz);
        const mk_alpha_wwwl& aw = P.alpha_wwwl[label];
        if (aw.alpha > 0) {
            T t = T(32768) * tex3_linear<T, Q8>(vol, p[0], p[1], p[2]);
            t = (T)((double)((t - T(aw.wl)) / T(aw.ww)) + 0.5);
            if (aAcc < t) aAcc = t;
        }
        acc += step;
    }
    T sum[4];
    if (aAcc <= T(0)) for (int c = 0; c < 4; ++c) sum[c] = T(P.background[c]);
    else for (int c = 0; c < 4; ++c) sum[c] = aAcc;
    write_rgb<T>(out, sum);
}

// d_renderSurface (core/kernel.cu:366-469).
template <typename T, bool Q8>
static void surface_pixel(const mk_render_params& P, const Volume& vol, int x, int y, uint8_t* out) {
    Ray<T> ray = make_ray<T>(P, x, y);
    T dirL[3];
    light_dir<T>(P, dirL);
    T sv[3] = {T(P.spacing_voxel[0]), T(P.spacing_voxel[1]), T(P.spacing_voxel[2])};
    const double scale = (double)P.scale;
    const float step = 1.0f / (float)vol.nz * P.step_scale;
    T aAcc = T(0);
    float acc = 0.0f;
    while ((double)acc < 1.732) {
        T p[3];
        ray.pos(acc, scale, p);
        int ix, iy, iz;
        if (!in_voi<T>(P.voi, vol, p, ix, iy, iz)) { acc += step; continue; }
        T t = T(32768) * tex3_linear<T, Q8>(vol, p[0], p[1], p[2]);
        if (t <= T(400)) { acc += step; continue; }
        T N[3];
        gradient<T, Q8>(vol, p, sv, ray.invz, N);
        aAcc = N[0] * dirL[0] + N[1] * dirL[1] + N[2] * dirL[2];
        if (aAcc > T(0)) break;
        acc += step;
    }
    T sum[4];
    if (aAcc <= T(0)) {
        for (int c = 0; c < 4; ++c) sum[c] = T(P.background[c]);
    } else {
        T g = (T)((double)aAcc + 0.05);  // NaN stays NaN and saturates to 0
        sum[0] = sum[1] = sum[2] = g; sum[3] = T(1);
    }
    write_rgb<T>(out, sum);
}

// d_renderPlane_{MIP,MinIP,Average} (core/kernel.cu:570-793) for one pixel.
template <typename T, bool Q8>
static int16_t plane_pixel(const mk_plane_params& P, const Volume& vol, int x, int y) {
    const T ps = T(P.pixel_spacing);
    const T half = T(P.half_num);
    const T W = T(vol.nx), H = T(vol.ny), D = T(vol.nz);
    short best = P.mpr_type == MK_MPR_MINIP ? (short)32767 : (short)-32768;
    double fsum = 0;
    for (T t = -half; t <= half; t += T(1)) {
        T fl = t * ps;
        T px = T(P.pt_left_top[0]) + fl * T(P.dir_n[0]);
        T py = T(P.pt_left_top[1]) + fl * T(P.dir_n[1]);
        T pz = T(P.pt_left_top[2]) + fl * T(P.dir_n[2]);
        T fx = (px + T(x) * ps * T(P.dir_h[0]) + T(y) * ps * T(P.dir_v[0])) / (T(P.spacing[0]) * W);
        T fy = (py + T(x) * ps * T(P.dir_h[1]) + T(y) * ps * T(P.dir_v[1])) / (T(P.spacing[1]) * H);
        T fz = (pz + T(x) * ps * T(P.dir_h[2]) + T(y) * ps * T(P.dir_v[2])) / (T(P.spacing[2]) * D);
        if (!P.invert_z) fz = (T)(1.0 - (double)fz);
        bool in = fx >= T(0) && fx <= T(1) && fy >= T(0) && fy <= T(1) && fz >= T(0) && fz <= T(1);
        T s = T(0);
        if (in) s = T(32768) * (P.sampling == MK_SAMPLE_NEAREST ? tex3_nearest<T>(vol, fx, fy, fz)
                                                                 : tex3_linear<T, Q8>(vol, fx, fy, fz));
        if (P.mpr_type == MK_MPR_AVERAGE) {
            fsum += in ? (double)s : -32768.0;
        } else {
            short val = in ? (short)s : (short)-32768;
            if (P.mpr_type == MK_MPR_MIP) best = best > val ? best : val;
            else best = best < val ? best : val;
        }
    }
    if (P.mpr_type == MK_MPR_AVERAGE) return (short)(fsum / (double)(T(2) * half + T(1)));
    return best;
}

// d_renderCPR (core/kernel.cu:839-873) for one pixel.
template <typename T, bool Q8>
static int16_t cpr_pixel(const double* pts, const double* dirs, int invz, int sampling, const Volume& vol, int x, int y) {
    const double* pt = pts + 3 * (size_t)y;
    const double* dr = dirs + 3 * (size_t)y;
    T fx = (T)((pt[0] + x * dr[0]) / (double)vol.nx);
    T fy = (T)((pt[1] + x * dr[1]) / (double)vol.ny);
    T fz = (T)((pt[2] + x * dr[2]) / (double)vol.nz);
    if (!invz) fz = (T)(1.0 - (double)fz);
    if (fx < T(0) || fx > T(1) || fy < T(0) || fy > T(1) || fz < T(0) || fz > T(1)) return -2048;
    T s = T(32768) * (sampling == MK_SAMPLE_NEAREST ? tex3_nearest<T>(vol, fx, fy, fz) : tex3_linear<T, Q8>(vol, fx, fy, fz));
    return (short)s;
}

template <typename T, bool Q8>
static void render_all(const mk_render_params& P, const Volume& vol, const Luts& L, uint8_t* out, int y0, int y1) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < P.width; ++x) {
            uint8_t* o = out + 3 * ((size_t)(y - y0) * P.width + x);
            if (P.render_type == MK_RENDER_VR) vr_pixel<T, Q8>(P, vol, L, x, y, o);
            else if (P.render_type == MK_RENDER_MIP) mip_pixel<T, Q8>(P, vol, x, y, o);
            else surface_pixel<T, Q8>(P, vol, x, y, o);
        }
    }
}

template <typename T, bool Q8>
static void plane_all(const mk_plane_params& P, const Volume& vol, int16_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < P.height; ++y)
        for (int x = 0; x < P.width; ++x) out[(size_t)y * P.width + x] = plane_pixel<T, Q8>(P, vol, x, y);
}

template <typename T, bool Q8>
static void cpr_all(int w, int h, const double* pts, const double* dirs, int invz, int sampling, const Volume& vol, int16_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) out[(size_t)y * w + x] = cpr_pixel<T, Q8>(pts, dirs, invz, sampling, vol, x, y);
}

// Fitting rule of mk_tf_forward in double (kernel.cu compositing on the
// float lattice at a fixed step, every sample with alpha > 0 composited, no
// early exit, alpha = min(1, s * LUT_alpha) with LUT entries in double, label
// 0 window, float RGB sum without saturation).
static void fit_pixel(const mk_render_params& P, const Volume& vol, const double* lut, int n, double s, int x, int y,
                      double out[3]) {
    Ray<double> ray = make_ray<double>(P, x, y);
    double dirL[3];
    light_dir<double>(P, dirL);
    double sv[3] = {P.spacing_voxel[0], P.spacing_voxel[1], P.spacing_voxel[2]};
    const float L1 = 1.0f / (float)vol.nz * P.step_scale;
    const double wl = P.alpha_wwwl[0].wl, ww = P.alpha_wwwl[0].ww;
    double C[3] = {0, 0, 0}, T = 1.0;
    float acc = 0.0f;
    while ((double)acc < 1.732) {
        double p[3];
        ray.pos(acc, (double)P.scale, p);
        acc += L1;
        int ix, iy, iz;
        if (!in_voi<double>(P.voi, vol, p, ix, iy, iz)) continue;
        double t = 32768.0 * tex3_linear<double, false>(vol, p[0], p[1], p[2]);
        t = (t - wl) / ww + 0.5;
        if (t > 1.0) t = 1.0;
        if (!(t >= -1.0)) t = -1.0;
        int i0, i1;
        double f;
        split<double, false>(t, n, i0, i1, f);
        double c[4];
        for (int k = 0; k < 4; ++k) c[k] = (1.0 - f) * lut[4 * i0 + k] + f * lut[4 * i1 + k];
        double a = s * c[3];
        a = a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a);
        if (!(a > 0.0)) continue;
        double N[3];
        gradient<double, false>(vol, p, sv, P.invert_z, N);
        double d = N[0] * dirL[0] + N[1] * dirL[1] + N[2] * dirL[2];
        double k = 0.35;
        if (d > 0.0) k += d * 0.8 + 0.16 * std::pow(d, 8.0);
        for (int ch = 0; ch < 3; ++ch) C[ch] += T * a * k * c[ch];
        T *= 1.0 - a;
    }
    out[0] = C[0]; out[1] = C[1]; out[2] = C[2];
}

// Marching cubes with the triangle-list contract of mk_marching_cubes: cubes
// x fastest, table order, vertices interpolated from the lower corner of
// each edge in float. Returns the total; writes min(total, capacity).
static uint64_t marching_cubes(const Volume& vol, const int* box, float iso, float* out, uint64_t capacity) {
    int8_t tri[256][16];
    uint8_t cnt[256];
    mk_mc_expand(tri, cnt);
    uint64_t t = 0;
    for (int k = box[2]; k < box[5] - 1; ++k)
        for (int j = box[1]; j < box[4] - 1; ++j)
            for (int i = box[0]; i < box[3] - 1; ++i) {
                float f[8];
                int c = 0;
                for (int n = 0; n < 8; ++n) {
                    f[n] = (float)vol.at(i + kMkMcCorner[n][0], j + kMkMcCorner[n][1], k + kMkMcCorner[n][2]);
                    if (f[n] < iso) c |= 1 << n;
                }
                for (int m = 0; m < cnt[c]; ++m, ++t) {
                    if (t >= capacity) continue;
                    for (int v = 0; v < 3; ++v) {
                        int e = tri[c][3 * m + v];
                        int a = kMkMcEdge[e][0], b = kMkMcEdge[e][1];
                        int sa = kMkMcCorner[a][0] + kMkMcCorner[a][1] + kMkMcCorner[a][2];
                        int sb = kMkMcCorner[b][0] + kMkMcCorner[b][1] + kMkMcCorner[b][2];
                        int lo = sa > sb ? b : a, hi = sa > sb ? a : b;
                        float tt = (iso - f[lo]) / (f[hi] - f[lo]);
                        float p[3] = {(float)(i + kMkMcCorner[lo][0]), (float)(j + kMkMcCorner[lo][1]),
                                      (float)(k + kMkMcCorner[lo][2])};
                        int axis = kMkMcCorner[lo][0] != kMkMcCorner[hi][0] ? 0 : (kMkMcCorner[lo][1] != kMkMcCorner[hi][1] ? 1 : 2);
                        p[axis] = p[axis] + tt;
                        for (int d = 0; d < 3; ++d) out[9 * t + 3 * v + d] = p[d];
                    }
                }
            }
    return t;
}

}  // namespace

extern "C" {

uint64_t mko_marching_cubes(const int16_t* v, int nx, int ny, int nz, const int* box, float iso, float* out,
                            uint64_t capacity) {
    Volume vol{v, nullptr, nx, ny, nz, nullptr};
    return marching_cubes(vol, box, iso, out, capacity);
}

// Fitting-rule image in double: out is height * width * 3.
int mko_fit_render(const mk_render_params* P, const int16_t* v, int nx, int ny, int nz, const double* lut, int n,
                   double s, double* out) {
    if (!P || !v || !lut || !out || n < 2) return 1;
    Volume vol{v, nullptr, nx, ny, nz, nullptr};
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < P->height; ++y)
        for (int x = 0; x < P->width; ++x) fit_pixel(*P, vol, lut, n, s, x, y, out + 3 * ((size_t)y * P->width + x));
    return 0;
}

// Rows [y0, y1) of the fitting-rule image in double: out is (y1 - y0) * width * 3.
int mko_fit_render_rows(const mk_render_params* P, const int16_t* v, int nx, int ny, int nz, const double* lut, int n,
                        double s, int y0, int y1, double* out) {
    if (!P || !v || !lut || !out || n < 2 || y0 < 0 || y1 > P->height || y0 > y1) return 1;
    Volume vol{v, nullptr, nx, ny, nz, nullptr};
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < P->width; ++x)
            fit_pixel(*P, vol, lut, n, s, x, y, out + 3 * ((size_t)(y - y0) * P->width + x));
    return 0;
}

// AnnotationUtils::Textout2Image (core/AnnotationUtils.cpp:419-464) and
// GetSize (466-475) after SetFontSize(font), over a glyph table instead of
// the reference's compiled-in masks: for glyph k = font * 256 + byte, width
// gw[k] and height gh[k] (-1 for a byte without a glyph, the reference's
// default Mask) and its bytes at atlas + goff[k]. img is RGB24, in place.
// Returns 0 when the reference returns false (nothing drawn), else 1.
int mko_textout(const char* text, int x, int y, int font, int format, const float* clr, const int* gw, const int* gh,
                const int64_t* goff, const uint8_t* atlas, uint8_t* img, int width, int height) {
    if (!img || width < 10 || height < 10) return 0;
    auto glyph = [&](unsigned char c, int& w, int& h, const uint8_t*& m) {
        w = -1; h = -1; m = nullptr;
        if (font < 0 || font > 2) return;   // FontSizeNotDefined: every mask is the default
        const int k = font * 256 + c;
        if (gw[k] > 0) { w = gw[k]; h = gh[k]; m = atlas + goff[k]; }
    };
    const size_t n = strlen(text);
    int xStart = x, yStart = y;
    if (format == -1) return 0;
    if (format == 2 || format == 1) {
        int w = 0;
        for (size_t i = 0; i < n; ++i) {   // GetSize
            int gw1, gh1;
            const uint8_t* m;
            glyph((unsigned char)text[i], gw1, gh1, m);
            w += gw1;
        }
        if (format == 2) xStart -= w / 2;
        else xStart -= w;
    }
    for (size_t i = 0; i < n; ++i) {
        int w, h;
        const uint8_t* m;
        glyph((unsigned char)text[i], w, h, m);
        for (int yy = yStart; yy < yStart + h; yy++) {
            for (int xx = xStart; xx < xStart + w; xx++) {
                if (xx < 0 || xx >= width || yy < 0 || yy >= height) continue;
                const uint8_t v = m[(yy - yStart) * w + xx - xStart];
                if (v > 0) {
                    // unsigned char * float in float, to unsigned char by
                    // truncation toward zero keeping the low 8 bits
                    for (int ch = 0; ch < 3; ++ch) {
                        float f = (float)v * clr[ch];
                        f = std::fmin(std::fmax(f, -2147483648.0f), 2147483520.0f);
                        img[(yy * (size_t)width + xx) * 3 + ch] = (uint8_t)(int)f;
                    }
                }
            }
        }
        xStart += w;
    }
    return 1;
}

// The expanded triangle table, for pinning against the reference's table.
void mko_mc_table(int8_t* tri, uint8_t* cnt) {
    mk_mc_expand((int8_t(*)[16])tri, cnt);
}

// precision: 0 = FP64 oracle, 1 = float CPU baseline. weights8: quantise
// filter weights to 8 fractional bits (float only).
int mko_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

int mko_render_rows(const mk_render_params* P, const int16_t* v, int nx, int ny, int nz, const uint8_t* mask,
                    const float* const* luts, const int* lut_len, int precision, int weights8, int y0, int y1,
                    uint8_t* out, uint8_t* touched);

int mko_render(const mk_render_params* P, const int16_t* v, int nx, int ny, int nz, const uint8_t* mask,
               const float* const* luts, const int* lut_len, int precision, int weights8, uint8_t* out,
               uint8_t* touched) {
    if (!P || !v || !out || nx <= 0 || ny <= 0 || nz <= 0 || !luts || !lut_len) return 1;
    return mko_render_rows(P, v, nx, ny, nz, mask, luts, lut_len, precision, weights8, 0, P->height, out, touched);
}

// Rows [y0, y1) of the frame P describes (rays unchanged), into out of
// (y1 - y0) * width * 3 bytes.
int mko_render_rows(const mk_render_params* P, const int16_t* v, int nx, int ny, int nz, const uint8_t* mask,
                    const float* const* luts, const int* lut_len, int precision, int weights8, int y0, int y1,
                    uint8_t* out, uint8_t* touched) {
    if (!P || !v || !out || nx <= 0 || ny <= 0 || nz <= 0 || !luts || !lut_len) return 1;
    if (y0 < 0 || y1 > P->height || y0 > y1) return 1;
    Volume vol{v, mask, nx, ny, nz, touched};
    Luts L{luts, lut_len};
    if (precision == 0) render_all<double, false>(*P, vol, L, out, y0, y1);
    else if (weights8) render_all<float, true>(*P, vol, L, out, y0, y1);
    else render_all<float, false>(*P, vol, L, out, y0, y1);
    return 0;
}

int mko_render_plane(const mk_plane_params* P, const int16_t* v, int nx, int ny, int nz, int precision,
                     int weights8, int16_t* out, uint8_t* touched) {
    if (!P || !v || !out || nx <= 0 || ny <= 0 || nz <= 0) return 1;
    Volume vol{v, nullptr, nx, ny, nz, touched};
    if (precision == 0) plane_all<double, false>(*P, vol, out);
    else if (weights8) plane_all<float, true>(*P, vol, out);
    else plane_all<float, false>(*P, vol, out);
    return 0;
}

int mko_render_cpr(int w, int h, const double* pts, const double* dirs, int invz, int sampling,
                   const int16_t* v, int nx, int ny, int nz, int precision, int weights8, int16_t* out,
                   uint8_t* touched) {
    if (!pts || !dirs || !v || !out || w <= 0 || h <= 0) return 1;
    Volume vol{v, nullptr, nx, ny, nz, touched};
    if (precision == 0) cpr_all<double, false>(w, h, pts, dirs, invz, sampling, vol, out);
    else if (weights8) cpr_all<float, true>(w, h, pts, dirs, invz, sampling, vol, out);
    else cpr_all<float, false>(w, h, pts, dirs, invz, sampling, vol, out);
    return 0;
}

// Single-sample probes used to pin the samplers against closed forms.
double mko_tex3_linear(const int16_t* v, int nx, int ny, int nz, double fx, double fy, double fz) {
    Volume vol{v, nullptr, nx, ny, nz, nullptr};
    return tex3_linear<double, false>(vol, fx, fy, fz);
}

int mko_tex1_linear(const float* lut, int n, double s, double* out4) {
    const float* luts[MK_NUM_LABELS] = {lut};
    int lens[MK_NUM_LABELS] = {n};
    Luts L{luts, lens};
    tex1_linear<double, false>(L, 0, s, out4);
    return 0;
}

}  // extern "C"
