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
//   * 16x16 pixel tiles of four 8x8 waves, mapped XCD-aware;
//   * the bricked layout's per-axis address terms of clamped voxel indices
//     read from tables in LDS (one instruction per axis for a sample's two
//     terms) instead of composed from bit fields;
//   * VR in two passes: k_vr marches every ray up to a trip budget (chosen
//     per frame) or until most of its waves have finished, and queues the
//     rays still marching with the state the reference's loop carries;
//     k_vr_tail finishes them eight lanes per ray, each lane evaluating one of
//     the ray's next samples at the current step and wavefront ballots finding
//     where the step machine backs up, changes step or exits. The frame is
//     the same, bit for bit, as a single pass's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mk_context.h"
#include "mk_render_dev.h"

mk_status mk_validate_render(mk_context* ctx, const mk_render_params* p) {
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

namespace {

constexpr int kMaxLdsBytes = 64 * 1024;
constexpr int kLutLdsEntries = 1024;  // stage LUTs in LDS up to 16 KiB
constexpr size_t kContHeader = 256;                  // queue counters
constexpr size_t kContMaxBytes = size_t(2) << 30;    // larger frames render in one pass
constexpr size_t kTailBlocksPerCu = 4;              // second-pass workgroups: one resident round (16 waves per CU)
constexpr float kLongSteps = 256.0f;                 // second-pass rays claimed first
constexpr int kTailGroup = 8;                        // second-pass lanes per ray
constexpr size_t kTailPercent = 95;                  // first-pass waves finished before the rest hand off
constexpr size_t kTermsLdsMax = 32 * 1024;           // axis term tables staged in LDS up to this
// Queue counters: 0 rays queued at the front, 1 rays claimed by the second
// pass, 2 rays queued at the back, 3 first-pass waves finished, 4 second-pass
// workgroups finished. All are zero when a frame starts.
constexpr int kQueueCounters = 5;
constexpr int kCounterGroupsDone = 4;

using namespace mkr;

// With the LUTs in LDS and the address terms composed, the first pass fits 96
// VGPRs: 5 waves per SIMD instead of 4 (1% faster on the body views). TAB:
// the samplers read the axis term tables staged after the accumulators; the
// compiler then takes ~99 VGPRs (4 waves per SIMD), which measured faster than
// holding it to 96 with two dwords spilled to scratch (whose write-backs
// doubled the pass's HBM writes).
template <int TW, bool MASK, bool LDS_LUT, bool TOUCH, bool TAB>
__global__ void __launch_bounds__(TW * TW, (LDS_LUT && !TOUCH && !TAB) ? 5 : 1) k_vr(RenderArgs A, AwTable T) {
    Smem s = stage_shared(A, T, LDS_LUT, A.nslot * TW * TW, TAB);
    int x, y;
    if (!tile_pixel<TW>(A, x, y)) return;
    vr_pixel<TW * TW, MASK, LDS_LUT, TOUCH, TAB>(A, s, x, y, threadIdx.x);
    // count the finished waves (rays queued or done): the first pass's last
    // waves hand off once most have finished
    if (!TOUCH && A.cont.cap > 0) {
        const uint64_t m = __ballot(1);
        if (__lane_id() == __ffsll((long long)m) - 1) atomicAdd(&A.cont.counters[3], 1u);
    }
}

// Second VR pass over the rays the first pass queued: 256 / kTailGroup rays in
// flight per workgroup (kTailGroup lanes per ray), claimed from the queue.
template <bool MASK, bool LDS_LUT, bool TAB>
__global__ void __launch_bounds__(256) k_vr_tail(RenderArgs A, AwTable T) {
    Smem s = stage_shared(A, T, LDS_LUT, (256 / kTailGroup) * (MASK ? A.nslot : 0), TAB);
    vr_tail<MASK, LDS_LUT, kTailGroup, TAB>(A, s, s.aobj + (threadIdx.x / kTailGroup) * (MASK ? A.nslot : 0));
    // The last workgroup to finish clears the queue counters for the next
    // frame: every other one has made its last counter access by then.
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t* c = A.cont.counters;
        if (atomicAdd(&c[kCounterGroupsDone], 1u) == gridDim.x - 1u)
            for (int i = 0; i < kQueueCounters; ++i) __hip_atomic_store(&c[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_zero_u32(uint32_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

template <int TW, bool MASK, bool TOUCH, bool TAB>
__global__ void __launch_bounds__(TW * TW) k_mip(RenderArgs A, AwTable T) {
    Smem s = stage_shared(A, T, false, 0, TAB);
    int x, y;
    if (!tile_pixel<TW>(A, x, y)) return;
    mip_pixel<MASK, TOUCH, TAB>(A, s, x, y);
}

template <int TW, bool TOUCH>
__global__ void __launch_bounds__(TW * TW) k_surface(RenderArgs A, AwTable T) {
    (void)stage_shared(A, T, false);
    int x, y;
    if (!tile_pixel<TW>(A, x, y)) return;
    surface_pixel<TOUCH>(A, x, y);
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

// The one-load brick entries (MK_BSKIP_*) from the brick and macro-cell
// occupancy and the brick labels.
__global__ void __launch_bounds__(256) k_brick_skip(const uint8_t* __restrict__ occ, const uint8_t* __restrict__ occ2,
                                                     const uint8_t* __restrict__ blabel, int nbx, int nby, int nmx,
                                                     int nmy, uint32_t nb, uint16_t* __restrict__ bskip) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const uint32_t bx = b % (uint32_t)nbx, by = (b / (uint32_t)nbx) % (uint32_t)nby, bz = b / ((uint32_t)nbx * nby);
    const uint32_t m = ((bz >> 2) * (uint32_t)nmy + (by >> 2)) * (uint32_t)nmx + (bx >> 2);
    bskip[b] = (uint16_t)(blabel[b] | (occ[b] ? MK_BSKIP_OCC : 0u) | (occ2[m] ? MK_BSKIP_MACRO : 0u));
}

mk_status launch_render(mk_context* ctx, const mk_render_params* p, uint8_t* dev_rgb, void* stream,
                        uint8_t* touch_v, uint8_t* touch_m) {
    mk_status st = mk_validate_render(ctx, p);
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
    st = mk_ensure_skip_tables(ctx, p->render_type, p->alpha_wwwl);
    if (st) return st;
    const float l1 = 1.0f / (float)ctx->nz * p->step_scale;
    st = mk_ensure_steps(ctx, l1);
    if (st) return st;

    RenderArgs A;
    memset(&A, 0, sizeof(A));
    A.v = ctx->view();
    A.lut = ctx->d_lut;
    A.tab = ctx->lut_tab;
    const MkSkipTables& sk = ctx->skip[p->render_type];
    A.occ = sk.occ;
    A.mipb = sk.mipb;
    A.occ2 = sk.occ2;
    A.blabel = ctx->stats.blabel;
    A.bskip = sk.bskip;
    A.mipb2 = sk.mipb2;
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
    A.rl1 = 1.0f / l1;
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

    // VR keeps a per-label accumulator per thread in LDS: with many labels
    // its tiles shrink to 8x8 (the MIP and Surface kernels keep none and
    // always take 16x16 tiles)
    const size_t acc_slots = p->render_type == MK_RENDER_VR ? (size_t)A.nslot : 0;
    int tw = 16;
    size_t lds = kLabelLdsBytes + (size_t)A.nlut_lds * sizeof(float4) + acc_slots * 256 * sizeof(float);
    if (lds > kMaxLdsBytes) {
        tw = 8;
        lds = kLabelLdsBytes + (size_t)A.nlut_lds * sizeof(float4) + acc_slots * 64 * sizeof(float);
        if (lds > kMaxLdsBytes) return mk_fail(ctx, MK_ERR_ARG, "too many labels for the LDS budget");
    }
    // the VR and MIP samplers take their address terms from LDS tables when
    // those fit (the footprint variant composes them)
    const size_t terms_bytes = sizeof(uint32_t) * (size_t)(axis_term_len(ctx->nx) + axis_term_len(ctx->ny) +
                                                           axis_term_len(ctx->nz));
    const bool tab = p->render_type != MK_RENDER_SURFACE && terms_bytes <= kTermsLdsMax &&
                     lds + terms_bytes <= kMaxLdsBytes && touch_v == nullptr;
    if (tab) lds += terms_bytes;
    const uint32_t tiles = (uint32_t)(((p->width + tw - 1) / tw) * ((p->height + tw - 1) / tw));
    hipStream_t s = (hipStream_t)stream;
    // the skip tables were (re)built on the context stream
    st = mk_order_after_context(ctx, s);
    if (st) return st;

    // VR in two passes: rays still marching after the trip budget are queued
    // and finished eight lanes per ray. The queue has a slot per pixel.
    const bool touch = touch_v != nullptr;
    const size_t npix = (size_t)p->width * (size_t)p->height;
    const int cslots = mask ? std::max(A.nslot, 1) : 1;
    const size_t cont_bytes = kContHeader + npix * (sizeof(float4) + 6 * sizeof(float) + (size_t)cslots * sizeof(float));
    int budget = ctx->vr_budget;
    if (ctx->n_cu <= 0) MK_HIP(ctx, hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    if (budget == kVrTripBudgetAuto) {
        const size_t waves = (size_t)tiles * (size_t)(tw * tw / 64);
        const size_t round = (size_t)ctx->n_cu * kVrResidentWavesPerCu;
        budget = waves <= round ? kVrTripBudgetSmall : (2 * waves <= 3 * round ? kVrTripBudgetMid : kVrTripBudgetLarge);
    }
    const bool two_pass = p->render_type == MK_RENDER_VR && !touch && budget > 0 && cont_bytes <= kContMaxBytes;
    if (two_pass) {
        // a larger queue replaces the buffer, after every reader of the
        // context (the last second pass among them) is done with the old one
        if (ctx->cont_bytes < cont_bytes) ctx->cont_clean = false;
        st = mk_ensure_buffer(ctx, &ctx->d_cont, &ctx->cont_bytes, cont_bytes);
        if (st) return st;
        // renders share the queue: one on another stream waits for the last
        // (one on the context's stream is already ordered by
        // mk_order_after_context, one on a caller's stream by that stream's
        // reader event, recorded after its second pass)
        if (ctx->cont_used && ctx->cont_stream != s && ctx->cont_stream != ctx->stream) {
            st = mk_order_after_reader(ctx, ctx->cont_stream, s);
            if (st) return st;
        }
        uint8_t* q = (uint8_t*)ctx->d_cont;
        VrCont& C = A.cont;
        C.counters = (uint32_t*)q;
        q += kContHeader;
        C.sum = (float4*)q;
        q += npix * sizeof(float4);
        C.pix = (int*)q;
        q += npix * sizeof(int);
        C.acc = (float*)q;
        q += npix * sizeof(float);
        C.step = (float*)q;
        q += npix * sizeof(float);
        C.apre = (float*)q;
        q += npix * sizeof(float);
        C.aacc = (float*)q;
        q += npix * sizeof(float);
        C.it = (int*)q;
        q += npix * sizeof(int);
        C.aobj = (float*)q;
        C.cap = (int)npix;
        C.budget = budget;
        C.long_steps = kLongSteps;
        // waves are 8x8 pixel blocks; the ones holding no pixel of the image
        // return at once and never count themselves
        const size_t live_waves = (size_t)((p->width + 7) / 8) * (size_t)((p->height + 7) / 8);
        C.tail_waves = (uint32_t)(live_waves * kTailPercent / 100);
        // each second pass leaves the counters at zero; a new buffer, or a
        // frame whose second pass was not enqueued, needs them cleared
        if (!ctx->cont_clean) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, C.counters, kQueueCounters);
        ctx->cont_clean = false;
    }
    dim3 grid(tiles), block(tw * tw);
#define MK_LAUNCH(K) hipLaunchKernelGGL(K, grid, block, lds, s, A, T)
#define MK_VR(TOUCH, TAB)                                                                         \
    if (tw == 16) {                                                                               \
        if (mask) { if (lds_lut) MK_LAUNCH((k_vr<16, true, true, TOUCH, TAB>)); else MK_LAUNCH((k_vr<16, true, false, TOUCH, TAB>)); } \
        else { if (lds_lut) MK_LAUNCH((k_vr<16, false, true, TOUCH, TAB>)); else MK_LAUNCH((k_vr<16, false, false, TOUCH, TAB>)); } \
    } else {                                                                                      \
        if (lds_lut) MK_LAUNCH((k_vr<8, true, true, TOUCH, TAB>)); else MK_LAUNCH((k_vr<8, true, false, TOUCH, TAB>)); \
    }
    if (p->render_type == MK_RENDER_VR) {
        if (touch) { MK_VR(true, false) } else if (tab) { MK_VR(false, true) } else { MK_VR(false, false) }
        if (two_pass) {
            const size_t groups = 256 / kTailGroup;
            const size_t lds2 = kLabelLdsBytes + (size_t)A.nlut_lds * sizeof(float4) + groups * (size_t)cslots * sizeof(float) +
                                (tab ? terms_bytes : 0);
            const size_t tail_blocks = (size_t)ctx->n_cu * kTailBlocksPerCu;
            const dim3 grid2((uint32_t)std::min<size_t>((npix + groups - 1) / groups, tail_blocks)), block2(256);
#define MK_TAIL(K) hipLaunchKernelGGL(K, grid2, block2, lds2, s, A, T)
            if (tab) {
                if (mask) { if (lds_lut) MK_TAIL((k_vr_tail<true, true, true>)); else MK_TAIL((k_vr_tail<true, false, true>)); }
                else { if (lds_lut) MK_TAIL((k_vr_tail<false, true, true>)); else MK_TAIL((k_vr_tail<false, false, true>)); }
            } else {
                if (mask) { if (lds_lut) MK_TAIL((k_vr_tail<true, true, false>)); else MK_TAIL((k_vr_tail<true, false, false>)); }
                else { if (lds_lut) MK_TAIL((k_vr_tail<false, true, false>)); else MK_TAIL((k_vr_tail<false, false, false>)); }
            }
#undef MK_TAIL
            MK_HIP(ctx, hipGetLastError());
            ctx->cont_clean = true;
            ctx->cont_stream = s;
            ctx->cont_used = true;
        }
    } else if (p->render_type == MK_RENDER_MIP) {
        if (touch) { if (mask) MK_LAUNCH((k_mip<16, true, true, false>)); else MK_LAUNCH((k_mip<16, false, true, false>)); }
        else if (tab) { if (mask) MK_LAUNCH((k_mip<16, true, false, true>)); else MK_LAUNCH((k_mip<16, false, false, true>)); }
        else { if (mask) MK_LAUNCH((k_mip<16, true, false, false>)); else MK_LAUNCH((k_mip<16, false, false, false>)); }
    } else {
        if (touch) MK_LAUNCH((k_surface<16, true>)); else MK_LAUNCH((k_surface<16, false>));
    }
#undef MK_VR
#undef MK_LAUNCH
    MK_HIP(ctx, hipGetLastError());
    return mk_note_reader(ctx, s);
}

}  // namespace

mk_status mk_ensure_skip_tables(mk_context* ctx, int render_type, const mk_alpha_wwwl* aw) {
    MkSkipTables& sk = ctx->skip[render_type];
    if (sk.valid && sk.lut_version == ctx->lut_version && sk.vol_version == ctx->vol_version &&
        memcmp(sk.aw, aw, sizeof(sk.aw)) == 0)
        return MK_OK;
    // frames on callers' streams may still read this render type's tables
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    AwTable T;
    for (int l = 0; l < MK_NUM_LABELS; ++l) {
        T.alpha[l] = aw[l].alpha;
        T.ww[l] = aw[l].ww;
        T.wl[l] = aw[l].wl;
    }
    uint32_t nb = (uint32_t)ctx->nbricks();
    hipLaunchKernelGGL(k_occupancy, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, ctx->stats.vmin, ctx->stats.vmax,
                       ctx->stats.lab_lo, ctx->stats.lab_hi, nb, render_type, T, ctx->lut_tab, ctx->d_lutnz, sk.occ,
                       sk.mipb);
    MK_HIP(ctx, hipGetLastError());
    const int nmx = (ctx->nbx + 3) / 4, nmy = (ctx->nby + 3) / 4, nmz = (ctx->nbz + 3) / 4;
    const uint32_t nm = (uint32_t)(nmx * nmy * nmz);
    hipLaunchKernelGGL(k_macro, dim3((nm + 255) / 256), dim3(256), 0, ctx->stream, sk.occ, sk.mipb, ctx->nbx,
                       ctx->nby, ctx->nbz, nmx, nmy, nmz, render_type, sk.occ2, sk.mipb2);
    MK_HIP(ctx, hipGetLastError());
    if (render_type != MK_RENDER_MIP) {
        hipLaunchKernelGGL(k_brick_skip, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, sk.occ, sk.occ2,
                           ctx->stats.blabel, ctx->nbx, ctx->nby, nmx, nmy, nb, sk.bskip);
        MK_HIP(ctx, hipGetLastError());
    }
    sk.valid = true;
    sk.lut_version = ctx->lut_version;
    sk.vol_version = ctx->vol_version;
    memcpy(sk.aw, aw, sizeof(sk.aw));
    return MK_OK;
}

mk_status mk_ensure_steps(mk_context* ctx, float l1) {
    if (ctx->steps_l1 == l1 && ctx->d_steps) return MK_OK;
    // frames on callers' streams may still read the current lattice
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    std::vector<float> a;
    a.reserve((size_t)(1.75 / l1) + 8);
    float acc = 0.0f;
    a.push_back(acc);
    while ((double)acc < 1.74) {
        acc += l1;
        a.push_back(acc);
    }
    st = mk_ensure_buffer(ctx, (void**)&ctx->d_steps, &ctx->steps_cap, a.size() * sizeof(float));
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

extern "C" mk_status mk_set_vr_trip_budget(mk_context* ctx, int trips) {
    if (!ctx) return MK_ERR_ARG;
    if (trips < kVrTripBudgetAuto) return mk_fail(ctx, MK_ERR_ARG, "trip budget %d: give trips >= 0, or -1 for automatic", trips);
    ctx->vr_budget = trips;
    return MK_OK;
}

extern "C" mk_status mk_render_footprint(mk_context* ctx, const mk_render_params* p, uint8_t* dev_rgb,
                                         uint8_t* dev_touch_vol, uint8_t* dev_touch_mask, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev_touch_vol) return mk_fail(ctx, MK_ERR_ARG, "null footprint map");
    return launch_render(ctx, p, dev_rgb, stream, dev_touch_vol, ctx->d_mask ? dev_touch_mask : nullptr);
}

extern "C" mk_status mk_render(mk_context* ctx, const mk_render_params* p, uint8_t* host_rgb) {
    if (!ctx) return MK_ERR_ARG;
    mk_status st = mk_validate_render(ctx, p);
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
