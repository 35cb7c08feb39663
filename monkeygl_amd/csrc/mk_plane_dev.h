// mk_plane_dev.h -- device code of slab MPR and CPR, shared by the per-image
// kernels (mk_plane.hip) and the batched serving kernel (mk_batch.hip).
// Floating-point contraction is off in these functions so nearest sampling
// is bit-exact with the host restatement wherever they are inlined.
#pragma once

#include <hip/hip_runtime.h>

#include "mk_context.h"

namespace mkp {

struct PlaneArgs {
    MkVolView v;
    int16_t* out;
    int width, height;
    int invert_z;
    float sp[3];
    float dh[3], dv[3], dn[3];
    float lt[3];
    float ps;
    float half;
};

// One slab-MPR pixel (d_renderPlane_*, core/kernel.cu:570-793).
template <int MODE, bool NEAREST>
__device__ inline void plane_pixel(const PlaneArgs& A, int x, int y) {
#pragma clang fp contract(off)
    const float W = (float)A.v.nx, H = (float)A.v.ny, D = (float)A.v.nz;
    short best = MODE == MK_MPR_MINIP ? (short)32767 : (short)-32768;
    double fsum = 0.0;
    for (float tt = -A.half; tt <= A.half; tt += 1.0f) {
        float fl = tt * A.ps;
        float px = A.lt[0] + fl * A.dn[0];
        float py = A.lt[1] + fl * A.dn[1];
        float pz = A.lt[2] + fl * A.dn[2];
        float fx = (px + (float)x * A.ps * A.dh[0] + (float)y * A.ps * A.dv[0]) / (A.sp[0] * W);
        float fy = (py + (float)x * A.ps * A.dh[1] + (float)y * A.ps * A.dv[1]) / (A.sp[1] * H);
        float fz = (pz + (float)x * A.ps * A.dh[2] + (float)y * A.ps * A.dv[2]) / (A.sp[2] * D);
        if (!A.invert_z) fz = (float)(1.0 - (double)fz);
        const bool in = fx >= 0.f && fx <= 1.f && fy >= 0.f && fy <= 1.f && fz >= 0.f && fz <= 1.f;
        float s = 0.f;
        if (in) s = NEAREST ? mk_sample_nearest_hu(A.v, fx, fy, fz) : mk_sample_hu(A.v, fx, fy, fz);
        if (MODE == MK_MPR_AVERAGE) {
            fsum += in ? (double)s : -32768.0;
        } else {
            short val = in ? (short)s : (short)-32768;
            if (MODE == MK_MPR_MIP) best = best > val ? best : val;
            else best = best < val ? best : val;
        }
    }
    short r = best;
    if (MODE == MK_MPR_AVERAGE) r = (short)(fsum / (double)(2.0f * A.half + 1.0f));
    A.out[(size_t)y * A.width + x] = r;
}

struct CprArgs {
    MkVolView v;
    const double* pts;
    const double* dirs;
    int16_t* out;
    int width, height;
    int invert_z;
};

// One CPR pixel (d_renderCPR, core/kernel.cu:839-873).
template <bool NEAREST>
__device__ inline void cpr_pixel(const CprArgs& A, int x, int y) {
#pragma clang fp contract(off)
    const double* pt = A.pts + 3 * (size_t)y;
    const double* dr = A.dirs + 3 * (size_t)y;
    float fx = (float)((pt[0] + x * dr[0]) / (double)A.v.nx);
    float fy = (float)((pt[1] + x * dr[1]) / (double)A.v.ny);
    float fz = (float)((pt[2] + x * dr[2]) / (double)A.v.nz);
    if (!A.invert_z) fz = (float)(1.0 - (double)fz);
    short r;
    if (fx < 0.f || fx > 1.f || fy < 0.f || fy > 1.f || fz < 0.f || fz > 1.f) r = -2048;
    else r = (short)(NEAREST ? mk_sample_nearest_hu(A.v, fx, fy, fz) : mk_sample_hu(A.v, fx, fy, fz));
    A.out[(size_t)y * A.width + x] = r;
}

}  // namespace mkp
