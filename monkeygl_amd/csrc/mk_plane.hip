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
#include "mk_plane_dev.h"

namespace {

using namespace mkp;

template <int MODE, bool NEAREST>
__global__ void __launch_bounds__(256) k_plane(PlaneArgs A) {
    const uint32_t tiles_x = (A.width + 15) / 16;
    const uint32_t tiles = tiles_x * ((A.height + 15) / 16);
    const uint32_t t = mk_xcd_tile(blockIdx.x, tiles);
    const int x = (int)(t % tiles_x) * 16 + (threadIdx.x & 15);
    const int y = (int)(t / tiles_x) * 16 + (threadIdx.x >> 4);
    if (x >= A.width || y >= A.height) return;
    plane_pixel<MODE, NEAREST>(A, x, y);
}

template <bool NEAREST>
__global__ void __launch_bounds__(256) k_cpr(CprArgs A) {
    const uint32_t tiles_x = (A.width + 15) / 16;
    const uint32_t tiles = tiles_x * ((A.height + 15) / 16);
    const uint32_t t = mk_xcd_tile(blockIdx.x, tiles);
    const int x = (int)(t % tiles_x) * 16 + (threadIdx.x & 15);
    const int y = (int)(t / tiles_x) * 16 + (threadIdx.x >> 4);
    if (x >= A.width || y >= A.height) return;
    cpr_pixel<NEAREST>(A, x, y);
}


}  // namespace

mk_status mk_validate_plane(mk_context* ctx, const mk_plane_params* p) {
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

mk_status mk_validate_cpr(mk_context* ctx, int width, int height, int sampling) {
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "CPR before any volume was uploaded");
    if (width <= 0 || height <= 0 || width > 65536 || height > 65536)
        return mk_fail(ctx, MK_ERR_ARG, "bad CPR size %dx%d", width, height);
    if (sampling != MK_SAMPLE_LINEAR && sampling != MK_SAMPLE_NEAREST)
        return mk_fail(ctx, MK_ERR_ARG, "unknown sampling %d", sampling);
    return MK_OK;
}

extern "C" mk_status mk_render_plane_async(mk_context* ctx, const mk_plane_params* p, int16_t* dev_out, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    mk_status st = mk_validate_plane(ctx, p);
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
    return mk_note_reader(ctx, s);
}

extern "C" mk_status mk_render_plane(mk_context* ctx, const mk_plane_params* p, int16_t* host_out) {
    if (!ctx) return MK_ERR_ARG;
    mk_status st = mk_validate_plane(ctx, p);
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
    mk_status vst = mk_validate_cpr(ctx, width, height, sampling);
    if (vst) return vst;
    if (!dev_points || !dev_dirs || !dev_out) return mk_fail(ctx, MK_ERR_ARG, "null CPR buffer");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    mk_status st = mk_order_after_context(ctx, s);
    if (st) return st;
    dim3 grid((uint32_t)(((width + 15) / 16) * ((height + 15) / 16))), block(256);
    CprArgs C;
    C.v = ctx->view();
    C.pts = dev_points;
    C.dirs = dev_dirs;
    C.out = dev_out;
    C.width = width;
    C.height = height;
    C.invert_z = invert_z != 0;
    if (sampling == MK_SAMPLE_NEAREST)
        hipLaunchKernelGGL(k_cpr<true>, grid, block, 0, s, C);
    else
        hipLaunchKernelGGL(k_cpr<false>, grid, block, 0, s, C);
    MK_HIP(ctx, hipGetLastError());
    return mk_note_reader(ctx, s);
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
