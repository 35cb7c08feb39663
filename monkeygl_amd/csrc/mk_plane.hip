// mk_plane.hip -- slab MPR and CPR reslicing on gfx950.
//
// Semantics follow d_renderPlane_MIP / _MinIP / _Average and d_renderCPR
// (core/kernel.cu:570-873): the same float plane geometry, in the same
// operation order (this file is compiled with -ffp-contract=off so nearest
// sampling is bit-exact with the host restatement), the same in-volume test
// and the same sentinels (-32768 for MPR, -2048 for CPR). Sampling reads the
// bricked int16 volume explicitly; MK_SAMPLE_NEAREST is a new mode.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "mk_context.h"

namespace {

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

template <int MODE, bool NEAREST>
__global__ void __launch_bounds__(256) k_plane(PlaneArgs A) {
    const uint32_t tiles_x = (A.width + 15) / 16;
    const uint32_t tiles = tiles_x * ((A.height + 15) / 16);
    const uint32_t t = mk_xcd_tile(blockIdx.x, tiles);
    const int x = (int)(t % tiles_x) * 16 + (threadIdx.x & 15);
    const int y = (int)(t / tiles_x) * 16 + (threadIdx.x >> 4);
    if (x >= A.width || y >= A.height) return;
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

template <bool NEAREST>
__global__ void __launch_bounds__(256) k_cpr(MkVolView v, const double* __restrict__ pts, const double* __restrict__ dirs,
                                             int width, int height, int invert_z, int16_t* __restrict__ out) {
    const uint32_t tiles_x = (width + 15) / 16;
    const uint32_t tiles = tiles_x * ((height + 15) / 16);
    const uint32_t t = mk_xcd_tile(blockIdx.x, tiles);
    const int x = (int)(t % tiles_x) * 16 + (threadIdx.x & 15);
    const int y = (int)(t / tiles_x) * 16 + (threadIdx.x >> 4);
    if (x >= width || y >= height) return;
    const double* pt = pts + 3 * (size_t)y;
    const double* dr = dirs + 3 * (size_t)y;
    float fx = (float)((pt[0] + x * dr[0]) / (double)v.nx);
    float fy = (float)((pt[1] + x * dr[1]) / (double)v.ny);
    float fz = (float)((pt[2] + x * dr[2]) / (double)v.nz);
    if (!invert_z) fz = (float)(1.0 - (double)fz);
    short r;
    if (fx < 0.f || fx > 1.f || fy < 0.f || fy > 1.f || fz < 0.f || fz > 1.f) r = -2048;
    else r = (short)(NEAREST ? mk_sample_nearest_hu(v, fx, fy, fz) : mk_sample_hu(v, fx, fy, fz));
    out[(size_t)y * width + x] = r;
}

mk_status validate_plane(mk_context* ctx, const mk_plane_params* p) {
    if (!p) return mk_fail(ctx, MK_ERR_ARG, "plane params are null");
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "plane render before any volume was uploaded");
    if (p->width <= 0 || p->height <= 0 || p->width > 65536 || p->height > 65536)
        return mk_fail(ctx, MK_ERR_ARG, "bad plane size %dx%d", p->width, p->height);
    if (p->mpr_type < MK_MPR_AVERAGE || p->mpr_type > MK_MPR_MINIP)
        return mk_fail(ctx, MK_ERR_ARG, "unknown MPR type %d", p->mpr_type);
    if (p->sampling != MK_SAMPLE_LINEAR && p->sampling != MK_SAMPLE_NEAREST)
        return mk_fail(ctx, MK_ERR_ARG, "unknown sampling %d", p->sampling);
    if (!std::isfinite(p->half_num) || p->half_num < 0.0f || p->half_num > 65536.0f)
        return mk_fail(ctx, MK_ERR_ARG, "bad half_num %g", (double)p->half_num);
    if (!(p->pixel_spacing > 0.0f) || !std::isfinite(p->pixel_spacing))
        return mk_fail(ctx, MK_ERR_ARG, "bad pixel spacing %g", (double)p->pixel_spacing);
    return MK_OK;
}

}  // namespace

extern "C" mk_status mk_render_plane_async(mk_context* ctx, const mk_plane_params* p, int16_t* dev_out, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    mk_status st = validate_plane(ctx, p);
    if (st) return st;
    if (!dev_out) return mk_fail(ctx, MK_ERR_ARG, "null output");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    PlaneArgs A;
    A.v = ctx->view();
    A.out = dev_out;
    A.width = p->width;
    A.height = p->height;
    A.invert_z = p->invert_z != 0;
    for (int i = 0; i < 3; ++i) {
        A.sp[i] = p->spacing[i];
        A.dh[i] = p->dir_h[i];
        A.dv[i] = p->dir_v[i];
        A.dn[i] = p->dir_n[i];
        A.lt[i] = p->pt_left_top[i];
    }
    A.ps = p->pixel_spacing;
    A.half = p->half_num;
    hipStream_t s = (hipStream_t)stream;
    st = mk_order_after_context(ctx, s);
    if (st) return st;
    dim3 grid((uint32_t)(((p->width + 15) / 16) * ((p->height + 15) / 16))), block(256);
    const bool nn = p->sampling == MK_SAMPLE_NEAREST;
    switch (p->mpr_type) {
        case MK_MPR_AVERAGE:
            if (nn) hipLaunchKernelGGL((k_plane<MK_MPR_AVERAGE, true>), grid, block, 0, s, A);
            else hipLaunchKernelGGL((k_plane<MK_MPR_AVERAGE, false>), grid, block, 0, s, A);
            break;
        case MK_MPR_MIP:
            if (nn) hipLaunchKernelGGL((k_plane<MK_MPR_MIP, true>), grid, block, 0, s, A);
            else hipLaunchKernelGGL((k_plane<MK_MPR_MIP, false>), grid, block, 0, s, A);
            break;
        default:
            if (nn) hipLaunchKernelGGL((k_plane<MK_MPR_MINIP, true>), grid, block, 0, s, A);
            else hipLaunchKernelGGL((k_plane<MK_MPR_MINIP, false>), grid, block, 0, s, A);
            break;
    }
    MK_HIP(ctx, hipGetLastError());
    return MK_OK;
}

extern "C" mk_status mk_render_plane(mk_context* ctx, const mk_plane_params* p, int16_t* host_out) {
    if (!ctx) return MK_ERR_ARG;
    mk_status st = validate_plane(ctx, p);
    if (st) return st;
    if (!host_out) return mk_fail(ctx, MK_ERR_ARG, "null output");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    size_t bytes = (size_t)p->width * p->height * sizeof(int16_t);
    st = mk_ensure_buffer(ctx, &ctx->d_stage, &ctx->stage_cap, bytes);
    if (st) return st;
    st = mk_render_plane_async(ctx, p, (int16_t*)ctx->d_stage, ctx->stream);
    if (st) return st;
    MK_HIP(ctx, hipMemcpyAsync(host_out, ctx->d_stage, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}

extern "C" mk_status mk_render_cpr_async(mk_context* ctx, int width, int height, const double* dev_points,
                                         const double* dev_dirs, int invert_z, int sampling, int16_t* dev_out,
                                         void* stream) {
    if (!ctx) return MK_ERR_ARG;
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "CPR before any volume was uploaded");
    if (width <= 0 || height <= 0 || width > 65536 || height > 65536)
        return mk_fail(ctx, MK_ERR_ARG, "bad CPR size %dx%d", width, height);
    if (!dev_points || !dev_dirs || !dev_out) return mk_fail(ctx, MK_ERR_ARG, "null CPR buffer");
    if (sampling != MK_SAMPLE_LINEAR && sampling != MK_SAMPLE_NEAREST)
        return mk_fail(ctx, MK_ERR_ARG, "unknown sampling %d", sampling);
    MK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    mk_status st = mk_order_after_context(ctx, s);
    if (st) return st;
    dim3 grid((uint32_t)(((width + 15) / 16) * ((height + 15) / 16))), block(256);
    if (sampling == MK_SAMPLE_NEAREST)
        hipLaunchKernelGGL(k_cpr<true>, grid, block, 0, s, ctx->view(), dev_points, dev_dirs, width, height,
                           invert_z != 0, dev_out);
    else
        hipLaunchKernelGGL(k_cpr<false>, grid, block, 0, s, ctx->view(), dev_points, dev_dirs, width, height,
                           invert_z != 0, dev_out);
    MK_HIP(ctx, hipGetLastError());
    return MK_OK;
}

extern "C" mk_status mk_render_cpr(mk_context* ctx, int width, int height, const double* points, const double* dirs,
                                   int invert_z, int sampling, int16_t* host_out) {
    if (!ctx) return MK_ERR_ARG;
    // core/kernel.cu:887-888: a null point list renders nothing
    if (!points || !dirs || !host_out) return mk_fail(ctx, MK_ERR_ARG, "null CPR buffer");
    if (width <= 0 || height <= 0 || width > 65536 || height > 65536)
        return mk_fail(ctx, MK_ERR_ARG, "bad CPR size %dx%d", width, height);
    MK_HIP(ctx, hipSetDevice(ctx->device));
    size_t geo = (size_t)height * 3 * sizeof(double);
    size_t bytes = (size_t)width * height * sizeof(int16_t);
    mk_status st = mk_ensure_buffer(ctx, &ctx->d_stage2, &ctx->stage2_cap, 2 * geo);
    if (st) return st;
    st = mk_ensure_buffer(ctx, &ctx->d_stage, &ctx->stage_cap, bytes);
    if (st) return st;
    double* d_pts = (double*)ctx->d_stage2;
    double* d_dirs = d_pts + 3 * (size_t)height;
    MK_HIP(ctx, hipMemcpyAsync(d_pts, points, geo, hipMemcpyHostToDevice, ctx->stream));
    MK_HIP(ctx, hipMemcpyAsync(d_dirs, dirs, geo, hipMemcpyHostToDevice, ctx->stream));
    st = mk_render_cpr_async(ctx, width, height, d_pts, d_dirs, invert_z, sampling, (int16_t*)ctx->d_stage, ctx->stream);
    if (st) return st;
    MK_HIP(ctx, hipMemcpyAsync(host_out, ctx->d_stage, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}
