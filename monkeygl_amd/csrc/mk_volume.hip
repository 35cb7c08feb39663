// mk_volume.hip -- context lifetime and the resident data of the render path:
// bricked int16 volume, bricked uint8 label mask, per-brick statistics and the
// transfer-function LUTs. Replaces the reference's CuDataManager
// (core/CuDataManager.cu:28-163, core/CuDataManager.cpp:34-88) and the device
// pick of DeviceInfo (core/DeviceInfo.cpp:28-70).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "mk_context.h"

static std::mutex g_err_mu;
static std::string g_create_err = "no error";

mk_status mk_fail(mk_context* ctx, mk_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) {
        ctx->err = buf;
    } else {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_create_err = buf;
    }
    return st;
}

mk_status mk_hip_fail(mk_context* ctx, hipError_t e, const char* what) {
    mk_status st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MK_ERR_NOMEM : MK_ERR_HIP;
    return mk_fail(ctx, st, "%s failed: %s", what, hipGetErrorString(e));
}

mk_status mk_ensure_buffer(mk_context* ctx, void** buf, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *buf) return MK_OK;
    if (*buf) {
        // work on callers' streams may still read the old buffer
        mk_status st = mk_wait_readers(ctx);
        if (st) return st;
        MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        MK_HIP(ctx, hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
    }
    if (bytes == 0) return MK_OK;
    MK_HIP(ctx, hipMalloc(buf, bytes));
    *cap = bytes;
    return MK_OK;
}

template <typename P> static void mk_free(P*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

extern "C" int mk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" mk_status mk_context_create(int device, mk_context** out) {
    if (!out) return mk_fail(nullptr, MK_ERR_ARG, "mk_context_create: out is null");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return mk_fail(nullptr, MK_ERR_HIP, "mk_context_create: no HIP device (%s)", hipGetErrorString(e));
    if (device < 0) {
        // DeviceInfo: the device with the largest global memory
        size_t best = 0;
        device = 0;
        for (int i = 0; i < n; ++i) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, i) == hipSuccess && prop.totalGlobalMem > best) {
                best = prop.totalGlobalMem;
                device = i;
            }
        }
    }
    if (device >= n) return mk_fail(nullptr, MK_ERR_ARG, "mk_context_create: device %d of %d", device, n);
    e = hipSetDevice(device);
    if (e != hipSuccess) return mk_fail(nullptr, MK_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    mk_context* ctx = new mk_context();
    ctx->device = device;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->order_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->mc_done, hipEventDisableTiming);
    if (e != hipSuccess) {
        mk_context_destroy(ctx);
        return mk_fail(nullptr, MK_ERR_HIP, "context stream/events: %s", hipGetErrorString(e));
    }
    // per-device set-up: the marching-cubes triangle table and the device
    // address of the allocation-free entry point's scratch
    mk_status st = mk_mc_device_init(device);
    if (st) {
        mk_context_destroy(ctx);
        return st;
    }
    *out = ctx;
    return MK_OK;
}

static void mk_free_volume(mk_context* ctx) {
    mk_free(ctx->d_vox);
    mk_free(ctx->d_mask);
    mk_free(ctx->stats.vmin);
    mk_free(ctx->stats.vmax);
    mk_free(ctx->stats.lab_lo);
    mk_free(ctx->stats.lab_hi);
    mk_free(ctx->stats.blabel);
    for (auto& t : ctx->skip) {
        mk_free(t.occ);
        mk_free(t.mipb);
        mk_free(t.occ2);
        mk_free(t.mipb2);
        mk_free(t.bskip);
        t.valid = false;
    }
    ctx->nx = ctx->ny = ctx->nz = 0;
}

extern "C" void mk_context_destroy(mk_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (auto& r : ctx->readers) {
        (void)hipEventSynchronize(r.done);
        (void)hipEventDestroy(r.done);
    }
    ctx->readers.clear();
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    mk_free_volume(ctx);
    mk_free(ctx->d_lut);
    mk_free(ctx->d_lutnz);
    mk_free(ctx->d_cont);
    if (ctx->mc_done) (void)hipEventDestroy(ctx->mc_done);
    if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
    mk_free(ctx->d_steps);
    if (ctx->d_mc_scratch) (void)hipFree(ctx->d_mc_scratch);
    if (ctx->d_mc_total) (void)hipFree(ctx->d_mc_total);
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    if (ctx->d_stage2) (void)hipFree(ctx->d_stage2);
    if (ctx->d_orient) (void)hipFree(ctx->d_orient);
    if (ctx->d_glyphs) (void)hipFree(ctx->d_glyphs);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* mk_last_error(const mk_context* ctx) {
    if (ctx) return ctx->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_create_err.c_str();
}

extern "C" void* mk_context_stream(mk_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

mk_status mk_order_after_context(mk_context* ctx, hipStream_t s) {
    if (s == ctx->stream) return MK_OK;
    // an idle context stream has nothing left to order after
    hipError_t q = hipStreamQuery(ctx->stream);
    if (q == hipSuccess) return MK_OK;
    if (q != hipErrorNotReady) return mk_hip_fail(ctx, q, "querying the context stream");
    // the wait takes the event's state at this call, so one event serves all
    hipError_t e = hipEventRecord(ctx->order_ev, ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, ctx->order_ev, 0);
    if (e != hipSuccess) return mk_hip_fail(ctx, e, "ordering after the context stream");
    return MK_OK;
}

// Streams tracked per context before the list is drained (synchronized and
// emptied): callers normally use a handful.
static constexpr size_t kMaxReaders = 64;

mk_status mk_note_reader(mk_context* ctx, hipStream_t s) {
    if (s == ctx->stream) return MK_OK;
    for (auto& r : ctx->readers)
        if (r.stream == s) {
            MK_HIP(ctx, hipEventRecord(r.done, s));
            return MK_OK;
        }
    if (ctx->readers.size() >= kMaxReaders) {
        for (auto& r : ctx->readers) {
            MK_HIP(ctx, hipEventSynchronize(r.done));
            (void)hipEventDestroy(r.done);
        }
        ctx->readers.clear();
    }
    mk_context::Reader r{s, nullptr};
    MK_HIP(ctx, hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
    hipError_t e = hipEventRecord(r.done, s);
    if (e != hipSuccess) {
        (void)hipEventDestroy(r.done);
        return mk_hip_fail(ctx, e, "recording a reader of the context");
    }
    ctx->readers.push_back(r);
    return MK_OK;
}

mk_status mk_order_after_reader(mk_context* ctx, hipStream_t r, hipStream_t s) {
    for (auto& e : ctx->readers)
        if (e.stream == r) {
            MK_HIP(ctx, hipStreamWaitEvent(s, e.done, 0));
            break;
        }
    return MK_OK;
}

mk_status mk_wait_readers(mk_context* ctx) {
    for (auto& r : ctx->readers) MK_HIP(ctx, hipStreamWaitEvent(ctx->stream, r.done, 0));
    return MK_OK;
}

extern "C" mk_status mk_synchronize(mk_context* ctx) {
    if (!ctx) return MK_ERR_ARG;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}

// ---- device memory for bindings ------------------------------------------

extern "C" mk_status mk_device_alloc(mk_context* ctx, uint64_t bytes, void** dev_out) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev_out) return mk_fail(ctx, MK_ERR_ARG, "mk_device_alloc: null out");
    *dev_out = nullptr;
    if (bytes == 0) return mk_fail(ctx, MK_ERR_ARG, "mk_device_alloc: zero bytes");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    MK_HIP(ctx, hipMalloc(dev_out, (size_t)bytes));
    return MK_OK;
}

extern "C" mk_status mk_device_free(mk_context* ctx, void* dev) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev) return MK_OK;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MK_HIP(ctx, hipFree(dev));
    return MK_OK;
}

extern "C" mk_status mk_copy_to_host(mk_context* ctx, void* host, const void* dev, uint64_t bytes) {
    if (!ctx) return MK_ERR_ARG;
    if (bytes == 0) return MK_OK;
    if (!host || !dev) return mk_fail(ctx, MK_ERR_ARG, "mk_copy_to_host: null pointer");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    MK_HIP(ctx, hipMemcpyAsync(host, dev, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}

extern "C" mk_status mk_copy_to_device(mk_context* ctx, void* dev, const void* host, uint64_t bytes) {
    if (!ctx) return MK_ERR_ARG;
    if (bytes == 0) return MK_OK;
    if (!host || !dev) return mk_fail(ctx, MK_ERR_ARG, "mk_copy_to_device: null pointer");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    MK_HIP(ctx, hipMemcpyAsync(dev, host, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    // the host buffer may go away at return
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}

// ---- bricking ------------------------------------------------------------

// One workgroup of 512 threads per brick: thread o writes voxel o of brick b,
// in the volume's order (mk_inbrick: 4x4x4 blocks) or, for the label mask
// (ROWS), in x-rows. Voxels past the volume edge are zero (the samplers clamp
// indices, so they are never read).
template <typename E, bool ROWS>
__global__ void __launch_bounds__(512) k_brickify(const E* __restrict__ lin, E* __restrict__ out, int nx, int ny, int nz,
                                                  int nbx, int nby) {
    uint32_t b = blockIdx.x;
    uint32_t o = threadIdx.x;
    int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    // inverse of mk_inbrick: offset bits x0 x1 y0 y1 z0 z1 x2 y2 z2
    const int ix = ROWS ? (int)(o & 7) : (int)((o & 3) | ((o >> 4) & 4));
    const int iy = ROWS ? (int)((o >> 3) & 7) : (int)(((o >> 2) & 3) | ((o >> 5) & 4));
    const int iz = ROWS ? (int)(o >> 6) : (int)(((o >> 4) & 3) | ((o >> 6) & 4));
    int x = bx * 8 + ix, y = by * 8 + iy, z = bz * 8 + iz;
    E v = 0;
    if (x < nx && y < ny && z < nz) v = lin[((size_t)z * ny + y) * nx + x];
    out[((size_t)b << 9) | o] = v;
}

// One wave per brick: min/max over the brick plus a one-voxel apron.
__global__ void __launch_bounds__(256) k_brick_minmax(MkVolView v, int16_t* __restrict__ vmin, int16_t* __restrict__ vmax,
                                                       uint32_t nbricks) {
    uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (b >= nbricks) return;
    int bx = b % v.nbx, by = (b / v.nbx) % v.nby, bz = b / (v.nbx * v.nby);
    int lo = 32767, hi = -32768;
    // 10x10x10 apron box, clamped to the volume
    for (int i = lane; i < 1000; i += 64) {
        int dx = i % 10, dy = (i / 10) % 10, dz = i / 100;
        int x = mk_clampi(bx * 8 - 1 + dx, 0, v.nx - 1);
        int y = mk_clampi(by * 8 - 1 + dy, 0, v.ny - 1);
        int z = mk_clampi(bz * 8 - 1 + dz, 0, v.nz - 1);
        int s = v.vox[mk_voxel_offset(v, x, y, z)];
        lo = min(lo, s);
        hi = max(hi, s);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    if (lane == 0) {
        vmin[b] = (int16_t)lo;
        vmax[b] = (int16_t)hi;
    }
}

// Label statistics. A wave handles a brick, each lane one 8-voxel x-row of it
// (lane = z * 8 + y within the brick, the row's bytes at lane * 8); nvalid is
// the number of the row's voxels inside the volume (0 for rows beyond it).
__device__ inline void mk_row_labels(uint64_t row, int nvalid, uint32_t& lo0, uint32_t& lo1, uint32_t& hi,
                                     uint32_t& bad) {
    for (int k = 0; k < nvalid; ++k) {
        const uint32_t l = (uint32_t)(row >> (8 * k)) & 0xFFu;
        if (l < 32) lo0 |= 1u << l;
        else if (l < 64) lo1 |= 1u << (l - 32);
        else if (l == 64) hi = 1;
        else bad = 1;
    }
}

__device__ inline int mk_row_valid(const MkVolView& v, uint32_t b, int lane, int& x0, int& y, int& z) {
    const int bx = b % v.nbx, by = (b / v.nbx) % v.nby, bz = b / (v.nbx * v.nby);
    x0 = bx * 8;
    y = by * 8 + (lane & 7);
    z = bz * 8 + (lane >> 3);
    return (y < v.ny && z < v.nz) ? min(8, v.nx - x0) : 0;
}

// Reduces one brick's row labels over the wave, writes the brick's entries,
// and ORs them into the wave's running union u[4].
__device__ inline void mk_brick_label_out(uint32_t b, uint32_t lo0, uint32_t lo1, uint32_t hi, uint32_t bad,
                                          uint64_t* lab_lo, uint8_t* lab_hi, uint8_t* blabel, uint32_t* u) {
    for (int off = 32; off > 0; off >>= 1) {
        lo0 |= __shfl_xor(lo0, off);
        lo1 |= __shfl_xor(lo1, off);
        hi |= __shfl_xor(hi, off);
        bad |= __shfl_xor(bad, off);
    }
    if ((threadIdx.x & 63) == 0) {
        const uint64_t lo = ((uint64_t)lo1 << 32) | lo0;
        lab_lo[b] = lo;
        lab_hi[b] = (uint8_t)hi;
        const int n = __popcll(lo) + (hi ? 1 : 0);
        blabel[b] = n == 1 ? (uint8_t)(hi ? 64 : __ffsll((long long)lo) - 1) : (uint8_t)0xFF;
    }
    u[0] |= lo0;
    u[1] |= lo1;
    u[2] |= hi;
    u[3] |= bad;
}

// The workgroup's union into flags with one atomic per word: flags[0] the
// labels 0..63, flags[2] bit 0 label 64, flags[3] != 0 a label above it.
__device__ inline void mk_label_union_out(const uint32_t* u, unsigned long long* flags) {
    __shared__ uint32_t s_u[4][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int i = 0; i < 4; ++i) s_u[wave][i] = u[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r[4] = {0, 0, 0, 0};
        for (int w = 0; w < 4; ++w)
            for (int i = 0; i < 4; ++i) r[i] |= s_u[w][i];
        const unsigned long long lo = ((unsigned long long)r[1] << 32) | r[0];
        if (lo) atomicOr(&flags[0], lo);
        if (r[2]) atomicOr(&flags[2], 1ull);
        if (r[3]) atomicOr(&flags[3], 1ull);
    }
}

// Label sets of every brick (grid-stride, one wave per brick).
__global__ void __launch_bounds__(256) k_brick_labels(MkVolView v, uint64_t* __restrict__ lab_lo, uint8_t* __restrict__ lab_hi,
                                                       uint8_t* __restrict__ blabel, uint32_t nbricks,
                                                       unsigned long long* __restrict__ flags) {
    const int lane = threadIdx.x & 63;
    uint32_t u[4] = {0, 0, 0, 0};
    for (uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6); b < nbricks; b += gridDim.x * 4) {
        int x0, y, z;
        const int nvalid = mk_row_valid(v, b, lane, x0, y, z);
        const uint64_t row = *(const uint64_t*)(v.mask + ((size_t)b << 9) + lane * 8);
        uint32_t lo0 = 0, lo1 = 0, hi = 0, bad = 0;
        mk_row_labels(row, nvalid, lo0, lo1, hi, bad);
        mk_brick_label_out(b, lo0, lo1, hi, bad, lab_lo, lab_hi, blabel, u);
    }
    mk_label_union_out(u, flags);
}

// Merges a binary object mask (src, x fastest, the volume's dims) into the
// bricked label mask in place, with the label statistics of the result.
// Voxels with src > 0 take label; with update, voxels with src == 0 holding
// label become 0 (VolumeInfo::AddNewObjectMask / UpdateObjectMask,
// core/VolumeInfo.cpp:94-133). Without an old mask (has_old == 0) the mask
// starts at 0. Voxels of the brick padding stay 0.
__global__ void __launch_bounds__(256) k_mask_merge(MkVolView v, const uint8_t* __restrict__ src, uint8_t* __restrict__ mask,
                                                     int has_old, uint32_t label, int update, uint64_t* __restrict__ lab_lo,
                                                     uint8_t* __restrict__ lab_hi, uint8_t* __restrict__ blabel,
                                                     uint32_t nbricks, unsigned long long* __restrict__ flags) {
    const int lane = threadIdx.x & 63;
    uint32_t u[4] = {0, 0, 0, 0};
    for (uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6); b < nbricks; b += gridDim.x * 4) {
        int x0, y, z;
        const int nvalid = mk_row_valid(v, b, lane, x0, y, z);
        uint64_t* rp = (uint64_t*)(mask + ((size_t)b << 9) + lane * 8);
        const uint64_t old = has_old ? *rp : 0ull;
        uint64_t nw = old;
        if (nvalid > 0) {
            const uint8_t* s = src + ((size_t)z * v.ny + y) * v.nx + x0;
            for (int k = 0; k < nvalid; ++k) {
                uint32_t m = (uint32_t)(old >> (8 * k)) & 0xFFu;
                if (s[k] > 0) m = label;
                else if (update && m == label) m = 0;
                nw = (nw & ~(0xFFull << (8 * k))) | ((uint64_t)m << (8 * k));
            }
        }
        *rp = nw;
        uint32_t lo0 = 0, lo1 = 0, hi = 0, bad = 0;
        mk_row_labels(nw, nvalid, lo0, lo1, hi, bad);
        mk_brick_label_out(b, lo0, lo1, hi, bad, lab_lo, lab_hi, blabel, u);
    }
    mk_label_union_out(u, flags);
}

// Axial slice z of the volume, -2048 where the mask holds 0 (when there is a
// mask): the image of HelloMonkey::GetOriginData_pngString
// (core/HelloMonkey.cpp:261-319).
__global__ void __launch_bounds__(256) k_origin_slice(MkVolView v, int z, int16_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)v.nx * (uint32_t)v.ny) return;
    const int x = (int)(i % (uint32_t)v.nx), y = (int)(i / (uint32_t)v.nx);
    int16_t val = v.vox[mk_voxel_offset(v, x, y, z)];
    if (v.mask && v.mask[mk_mask_offset(v, x, y, z)] == 0) val = -2048;
    out[i] = val;
}

// Bricked bytes back to the x-fastest layout; one workgroup per brick.
__global__ void __launch_bounds__(512) k_unbrickify_u8(const uint8_t* __restrict__ br, uint8_t* __restrict__ lin, int nx,
                                                       int ny, int nz, int nbx, int nby) {
    const uint32_t b = blockIdx.x, o = threadIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int x = bx * 8 + (o & 7), y = by * 8 + ((o >> 3) & 7), z = bz * 8 + (o >> 6);
    if (x < nx && y < ny && z < nz) lin[((size_t)z * ny + y) * nx + x] = br[((size_t)b << 9) | o];
}

// grid of the grid-stride label kernels
static uint32_t mk_label_grid(size_t nb) { return (uint32_t)std::min<size_t>((nb + 3) / 4, 16384); }

__global__ void k_fill_u64(uint64_t* p, uint64_t val, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}
__global__ void k_fill_u8(uint8_t* p, uint8_t val, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}

static mk_status mk_check_dims(mk_context* ctx, int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return mk_fail(ctx, MK_ERR_ARG, "volume dims must be positive (%d,%d,%d)", nx, ny, nz);
    if (nx > MK_MAX_DIM || ny > MK_MAX_DIM || nz > MK_MAX_DIM)
        return mk_fail(ctx, MK_ERR_ARG, "volume dims (%d,%d,%d) above %d", nx, ny, nz, MK_MAX_DIM);
    size_t nb = (size_t)((nx + 7) / 8) * ((ny + 7) / 8) * ((nz + 7) / 8);
    if (nb * 512 > 0x80000000ull) return mk_fail(ctx, MK_ERR_ARG, "volume too large: %zu padded voxels (limit 2^31)", nb * 512);
    return MK_OK;
}

static mk_status mk_volume_from_device(mk_context* ctx, const int16_t* d_lin, int nx, int ny, int nz) {
    // renders on callers' streams may still read the old volume
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    mk_free_volume(ctx);
    ctx->nx = nx; ctx->ny = ny; ctx->nz = nz;
    ctx->nbx = (nx + 7) / 8; ctx->nby = (ny + 7) / 8; ctx->nbz = (nz + 7) / 8;
    size_t nb = ctx->nbricks();
    MK_HIP(ctx, hipMalloc((void**)&ctx->d_vox, nb * 512 * sizeof(int16_t)));
    MK_HIP(ctx, hipMalloc((void**)&ctx->stats.vmin, nb * sizeof(int16_t)));
    MK_HIP(ctx, hipMalloc((void**)&ctx->stats.vmax, nb * sizeof(int16_t)));
    MK_HIP(ctx, hipMalloc((void**)&ctx->stats.lab_lo, nb * sizeof(uint64_t)));
    MK_HIP(ctx, hipMalloc((void**)&ctx->stats.lab_hi, nb * sizeof(uint8_t)));
    MK_HIP(ctx, hipMalloc((void**)&ctx->stats.blabel, nb * sizeof(uint8_t)));
    const size_t nm = (size_t)((ctx->nbx + 3) / 4) * ((ctx->nby + 3) / 4) * ((ctx->nbz + 3) / 4);
    for (auto& t : ctx->skip) {
        MK_HIP(ctx, hipMalloc((void**)&t.occ, nb * sizeof(uint8_t)));
        MK_HIP(ctx, hipMalloc((void**)&t.mipb, nb * sizeof(float)));
        MK_HIP(ctx, hipMalloc((void**)&t.occ2, nm * sizeof(uint8_t)));
        MK_HIP(ctx, hipMalloc((void**)&t.mipb2, nm * sizeof(float)));
        MK_HIP(ctx, hipMalloc((void**)&t.bskip, nb * sizeof(uint16_t)));
    }
    hipLaunchKernelGGL((k_brickify<int16_t, false>), dim3((uint32_t)nb), dim3(512), 0, ctx->stream, d_lin, ctx->d_vox, nx, ny, nz,
                       ctx->nbx, ctx->nby);
    MK_HIP(ctx, hipGetLastError());
    MkVolView v = ctx->view();
    hipLaunchKernelGGL(k_brick_minmax, dim3((uint32_t)((nb + 3) / 4)), dim3(256), 0, ctx->stream, v, ctx->stats.vmin,
                       ctx->stats.vmax, (uint32_t)nb);
    MK_HIP(ctx, hipGetLastError());
    // no mask: every brick holds label 0 only
    hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.lab_lo,
                       (uint64_t)1, (uint32_t)nb);
    hipLaunchKernelGGL(k_fill_u8, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.lab_hi,
                       (uint8_t)0, (uint32_t)nb);
    hipLaunchKernelGGL(k_fill_u8, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.blabel,
                       (uint8_t)0, (uint32_t)nb);
    MK_HIP(ctx, hipGetLastError());
    ctx->used_lo = 1;
    ctx->used_hi = false;
    ctx->vol_version++;
    for (auto& t : ctx->skip) t.valid = false;
    return MK_OK;
}

extern "C" mk_status mk_copy_volume_data(mk_context* ctx, const int16_t* host, int nx, int ny, int nz) {
    if (!ctx) return MK_ERR_ARG;
    if (!host) return mk_fail(ctx, MK_ERR_ARG, "mk_copy_volume_data: null data");
    mk_status st = mk_check_dims(ctx, nx, ny, nz);
    if (st) return st;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    size_t bytes = (size_t)nx * ny * nz * sizeof(int16_t);
    void* tmp = nullptr;
    MK_HIP(ctx, hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, ctx->stream);
    st = e == hipSuccess ? mk_volume_from_device(ctx, (const int16_t*)tmp, nx, ny, nz) : mk_hip_fail(ctx, e, "hipMemcpyAsync");
    e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    if (st) return st;
    if (e != hipSuccess) return mk_hip_fail(ctx, e, "hipStreamSynchronize");
    return MK_OK;
}

extern "C" mk_status mk_copy_volume_data_device(mk_context* ctx, const int16_t* dev, int nx, int ny, int nz) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev) return mk_fail(ctx, MK_ERR_ARG, "mk_copy_volume_data_device: null data");
    mk_status st = mk_check_dims(ctx, nx, ny, nz);
    if (st) return st;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    st = mk_volume_from_device(ctx, dev, nx, ny, nz);
    if (st) return st;
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MK_OK;
}

static mk_status mk_mask_from_device(mk_context* ctx, const uint8_t* d_lin, int nx, int ny, int nz) {
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "mask upload before any volume");
    if (nx != ctx->nx || ny != ctx->ny || nz != ctx->nz)
        return mk_fail(ctx, MK_ERR_ARG, "mask dims (%d,%d,%d) differ from volume (%d,%d,%d)", nx, ny, nz, ctx->nx, ctx->ny, ctx->nz);
    // the label statistics are rewritten below, the old mask freed
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    size_t nb = ctx->nbricks();
    uint8_t* d_mask = nullptr;
    MK_HIP(ctx, hipMalloc((void**)&d_mask, nb * 512));
    hipLaunchKernelGGL((k_brickify<uint8_t, true>), dim3((uint32_t)nb), dim3(512), 0, ctx->stream, d_lin, d_mask, nx, ny, nz,
                       ctx->nbx, ctx->nby);
    MK_HIP(ctx, hipGetLastError());
    unsigned long long* d_flags = nullptr;
    MK_HIP(ctx, hipMalloc((void**)&d_flags, 4 * sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, (uint64_t*)d_flags, (uint64_t)0, 4u);
    MkVolView v = ctx->view();
    v.mask = d_mask;
    hipLaunchKernelGGL(k_brick_labels, dim3(mk_label_grid(nb)), dim3(256), 0, ctx->stream, v, ctx->stats.lab_lo,
                       ctx->stats.lab_hi, ctx->stats.blabel, (uint32_t)nb, d_flags);
    MK_HIP(ctx, hipGetLastError());
    unsigned long long flags[4];
    MK_HIP(ctx, hipMemcpyAsync(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(d_flags);
    if (flags[3]) {
        (void)hipFree(d_mask);
        // restore the no-mask label statistics
        hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.lab_lo,
                           (uint64_t)1, (uint32_t)nb);
        hipLaunchKernelGGL(k_fill_u8, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.lab_hi,
                           (uint8_t)0, (uint32_t)nb);
        hipLaunchKernelGGL(k_fill_u8, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.blabel,
                           (uint8_t)0, (uint32_t)nb);
        if (ctx->d_mask) {
            // a previous mask stays active: recompute its statistics
            MkVolView v2 = ctx->view();
            unsigned long long* d_f2 = nullptr;
            MK_HIP(ctx, hipMalloc((void**)&d_f2, 4 * sizeof(unsigned long long)));
            hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, (uint64_t*)d_f2, (uint64_t)0, 4u);
            hipLaunchKernelGGL(k_brick_labels, dim3(mk_label_grid(nb)), dim3(256), 0, ctx->stream, v2,
                               ctx->stats.lab_lo, ctx->stats.lab_hi, ctx->stats.blabel, (uint32_t)nb, d_f2);
            MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipFree(d_f2);
        }
        MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return mk_fail(ctx, MK_ERR_RANGE, "mask holds a label above %d", MK_MAX_OBJECTS);
    }
    if (ctx->d_mask) (void)hipFree(ctx->d_mask);
    ctx->d_mask = d_mask;
    ctx->used_lo = flags[0];
    ctx->used_hi = flags[2] != 0;
    ctx->vol_version++;
    for (auto& t : ctx->skip) t.valid = false;
    return MK_OK;
}

extern "C" mk_status mk_copy_mask_data(mk_context* ctx, const uint8_t* host, int nx, int ny, int nz) {
    if (!ctx) return MK_ERR_ARG;
    if (!host) return mk_fail(ctx, MK_ERR_ARG, "mk_copy_mask_data: null data");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    if (nx <= 0 || ny <= 0 || nz <= 0) return mk_fail(ctx, MK_ERR_ARG, "mask dims must be positive");
    size_t bytes = (size_t)nx * ny * nz;
    void* tmp = nullptr;
    MK_HIP(ctx, hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, ctx->stream);
    mk_status st = e == hipSuccess ? mk_mask_from_device(ctx, (const uint8_t*)tmp, nx, ny, nz) : mk_hip_fail(ctx, e, "hipMemcpyAsync");
    e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    if (st) return st;
    if (e != hipSuccess) return mk_hip_fail(ctx, e, "hipStreamSynchronize");
    return MK_OK;
}

extern "C" mk_status mk_copy_mask_data_device(mk_context* ctx, const uint8_t* dev, int nx, int ny, int nz) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev) return mk_fail(ctx, MK_ERR_ARG, "mk_copy_mask_data_device: null data");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    return mk_mask_from_device(ctx, dev, nx, ny, nz);
}

extern "C" mk_status mk_clear_mask(mk_context* ctx) {
    if (!ctx) return MK_ERR_ARG;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->d_mask) return MK_OK;
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    mk_free(ctx->d_mask);
    size_t nb = ctx->nbricks();
    hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.lab_lo,
                       (uint64_t)1, (uint32_t)nb);
    hipLaunchKernelGGL(k_fill_u8, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.lab_hi,
                       (uint8_t)0, (uint32_t)nb);
    hipLaunchKernelGGL(k_fill_u8, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->stats.blabel,
                       (uint8_t)0, (uint32_t)nb);
    MK_HIP(ctx, hipGetLastError());
    ctx->used_lo = 1;
    ctx->used_hi = false;
    ctx->vol_version++;
    for (auto& t : ctx->skip) t.valid = false;
    return MK_OK;
}

// ---- incremental mask edits, slices -----------------------------------------

static mk_status mk_merge_from_device(mk_context* ctx, const uint8_t* d_lin, int nx, int ny, int nz, int label,
                                      int mode) {
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "mask merge before any volume");
    if (nx != ctx->nx || ny != ctx->ny || nz != ctx->nz)
        return mk_fail(ctx, MK_ERR_ARG, "mask dims (%d,%d,%d) differ from volume (%d,%d,%d)", nx, ny, nz, ctx->nx,
                       ctx->ny, ctx->nz);
    if (label < 0 || label > MK_MAX_OBJECTS)
        return mk_fail(ctx, MK_ERR_RANGE, "label %d outside 0..%d", label, MK_MAX_OBJECTS);
    if (mode != MK_MASK_ADD && mode != MK_MASK_UPDATE) return mk_fail(ctx, MK_ERR_ARG, "unknown merge mode %d", mode);
    if (mode == MK_MASK_UPDATE && !ctx->d_mask) return mk_fail(ctx, MK_ERR_STATE, "mask update without a mask");
    // the mask and its label statistics are rewritten in place
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    const size_t nb = ctx->nbricks();
    uint8_t* d_mask = ctx->d_mask;
    const int has_old = d_mask != nullptr;
    if (!has_old) MK_HIP(ctx, hipMalloc((void**)&d_mask, nb * 512));
    unsigned long long* d_flags = nullptr;
    hipError_t e = hipMalloc((void**)&d_flags, 4 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        if (!has_old) (void)hipFree(d_mask);
        return mk_hip_fail(ctx, e, "hipMalloc");
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, (uint64_t*)d_flags, (uint64_t)0, 4u);
    MkVolView v = ctx->view();
    hipLaunchKernelGGL(k_mask_merge, dim3(mk_label_grid(nb)), dim3(256), 0, ctx->stream, v, d_lin, d_mask, has_old,
                       (uint32_t)label, mode == MK_MASK_UPDATE ? 1 : 0, ctx->stats.lab_lo, ctx->stats.lab_hi,
                       ctx->stats.blabel, (uint32_t)nb, d_flags);
    unsigned long long flags[4] = {0, 0, 0, 0};
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_flags);
    if (e != hipSuccess) {
        if (!has_old) (void)hipFree(d_mask);
        return mk_hip_fail(ctx, e, "mask merge");
    }
    ctx->d_mask = d_mask;
    ctx->used_lo = flags[0];
    ctx->used_hi = flags[2] != 0;
    ctx->vol_version++;
    for (auto& t : ctx->skip) t.valid = false;
    return MK_OK;
}

extern "C" mk_status mk_merge_object_mask(mk_context* ctx, const uint8_t* host, int nx, int ny, int nz, int label,
                                          int mode) {
    if (!ctx) return MK_ERR_ARG;
    if (!host) return mk_fail(ctx, MK_ERR_ARG, "mk_merge_object_mask: null data");
    if (nx <= 0 || ny <= 0 || nz <= 0) return mk_fail(ctx, MK_ERR_ARG, "mask dims must be positive");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = (size_t)nx * ny * nz;
    void* tmp = nullptr;
    MK_HIP(ctx, hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, ctx->stream);
    mk_status st = e == hipSuccess ? mk_merge_from_device(ctx, (const uint8_t*)tmp, nx, ny, nz, label, mode)
                                   : mk_hip_fail(ctx, e, "hipMemcpyAsync");
    e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    if (st) return st;
    if (e != hipSuccess) return mk_hip_fail(ctx, e, "hipStreamSynchronize");
    return MK_OK;
}

extern "C" mk_status mk_merge_object_mask_device(mk_context* ctx, const uint8_t* dev, int nx, int ny, int nz,
                                                 int label, int mode) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev) return mk_fail(ctx, MK_ERR_ARG, "mk_merge_object_mask_device: null data");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    return mk_merge_from_device(ctx, dev, nx, ny, nz, label, mode);
}

extern "C" mk_status mk_get_mask_data(mk_context* ctx, uint8_t* host) {
    if (!ctx) return MK_ERR_ARG;
    if (!host) return mk_fail(ctx, MK_ERR_ARG, "mk_get_mask_data: null buffer");
    if (!ctx->d_mask) return mk_fail(ctx, MK_ERR_STATE, "no mask");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = (size_t)ctx->nx * ctx->ny * ctx->nz;
    void* tmp = nullptr;
    MK_HIP(ctx, hipMalloc(&tmp, bytes));
    hipLaunchKernelGGL(k_unbrickify_u8, dim3((uint32_t)ctx->nbricks()), dim3(512), 0, ctx->stream, ctx->d_mask,
                       (uint8_t*)tmp, ctx->nx, ctx->ny, ctx->nz, ctx->nbx, ctx->nby);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return mk_hip_fail(ctx, e, "mk_get_mask_data");
    return MK_OK;
}

extern "C" mk_status mk_origin_slice_async(mk_context* ctx, int z, int16_t* dev_out, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    if (!dev_out) return mk_fail(ctx, MK_ERR_ARG, "mk_origin_slice_async: null output");
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "no volume");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    z = z < 0 ? 0 : (z >= ctx->nz ? ctx->nz - 1 : z);
    hipStream_t s = (hipStream_t)stream;
    mk_status st = mk_order_after_context(ctx, s);
    if (st) return st;
    const uint32_t n = (uint32_t)ctx->nx * (uint32_t)ctx->ny;
    hipLaunchKernelGGL(k_origin_slice, dim3((n + 255) / 256), dim3(256), 0, s, ctx->view(), z, dev_out);
    MK_HIP(ctx, hipGetLastError());
    return mk_note_reader(ctx, s);
}

// ---- transfer functions ---------------------------------------------------

extern "C" mk_status mk_set_transfer_func(mk_context* ctx, int label, const float* rgba, int n) {
    if (!ctx) return MK_ERR_ARG;
    if (label < 0 || label > MK_MAX_OBJECTS) return mk_fail(ctx, MK_ERR_ARG, "label %d out of range", label);
    if (n < 0 || (n > 0 && !rgba)) return mk_fail(ctx, MK_ERR_ARG, "bad LUT (n=%d)", n);
    if (n > (1 << 20)) return mk_fail(ctx, MK_ERR_ARG, "LUT too long (%d entries)", n);
    ctx->h_lut[label].assign(rgba, rgba + 4 * (size_t)n);
    ctx->lut_dirty = true;
    return MK_OK;
}

mk_status mk_upload_luts(mk_context* ctx) {
    if (!ctx->lut_dirty && ctx->d_lut) return MK_OK;
    // frames on callers' streams may still read the current LUTs
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    size_t total = 0, total_nz = 0;
    for (int l = 0; l < MK_NUM_LABELS; ++l) {
        int len = (int)(ctx->h_lut[l].size() / 4);
        ctx->lut_tab.off[l] = (int)total;
        ctx->lut_tab.len[l] = len;
        total += len;
        total_nz += len + 1;
    }
    std::vector<float> packed(4 * (total ? total : 1), 0.0f);
    std::vector<int> nz(total_nz ? total_nz : 1, 0);
    for (int l = 0; l < MK_NUM_LABELS; ++l) {
        int len = ctx->lut_tab.len[l];
        if (len) memcpy(&packed[4 * (size_t)ctx->lut_tab.off[l]], ctx->h_lut[l].data(), 16 * (size_t)len);
        int* p = &nz[(size_t)ctx->lut_tab.off[l] + l];
        p[0] = 0;
        for (int i = 0; i < len; ++i) p[i + 1] = p[i] + (ctx->h_lut[l][4 * (size_t)i + 3] != 0.0f ? 1 : 0);
    }
    st = mk_ensure_buffer(ctx, (void**)&ctx->d_lut, &ctx->lut_cap, packed.size() * sizeof(float));
    if (st) return st;
    st = mk_ensure_buffer(ctx, (void**)&ctx->d_lutnz, &ctx->lutnz_cap, nz.size() * sizeof(int));
    if (st) return st;
    MK_HIP(ctx, hipMemcpyAsync(ctx->d_lut, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    MK_HIP(ctx, hipMemcpyAsync(ctx->d_lutnz, nz.data(), nz.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    // host vectors die at return: wait for the copies
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->lut_dirty = false;
    ctx->lut_version++;
    return MK_OK;
}
