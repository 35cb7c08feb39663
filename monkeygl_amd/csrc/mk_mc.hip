// mk_mc.hip -- marching-cubes iso-surface extraction on gfx950.
//
// The reference holds only the triangle table (core/MarchingCube.cpp:35-292);
// extraction is new. The triangle list is defined as:
//   * cubes are the unit cells with min corner (i, j, k) inside the box
//     [x0, x1-1) x [y0, y1-1) x [z0, z1-1), visited i fastest, then j, then k;
//   * corner values are the raw int16 voxels as float; corner c is inside
//     when value < iso and sets bit c of the case index;
//   * triangles of a cube follow the table order, vertices the edge order;
//   * a vertex on an edge interpolates from the corner with the smaller
//     coordinate along the edge: t = (iso - f_lo) / (f_hi - f_lo), coordinate
//     = lo + t, so an edge shared by neighbouring cubes yields the same bits.
// Three launches: per-block triangle counts, an exclusive scan of the block
// counts, and the emit pass that recomputes each cube and writes at its
// block's offset plus its rank in the block. This file is compiled with
// -ffp-contract=off so the vertices are bit-exact with the host restatement.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "mk_context.h"
#include "mk_mc_table.h"

namespace {

constexpr int kMcBlock = 256;

// Device-resident state shared by every context of a device, set up once per
// device by mk_mc_device_init (called from mk_context_create):
//   * the expanded triangle table: [256][16] edge indices, then [256] counts;
//   * `scratch`, the workspace of mk_marching_cubes_noalloc: per-block
//     triangle counts (uint32), per-block offsets (uint64, 256-byte aligned),
//     then the total (uint64). Its size bounds that entry point's box to
//     kMcNoAllocMaxBlocks blocks of kMcBlock cubes.
constexpr size_t kMcTableBytes = 256 * 16 + 256;
constexpr uint32_t kMcNoAllocMaxBlocks = 1u << 20;   // 2^28 cubes
constexpr size_t kMcScratchOffs = ((size_t)kMcNoAllocMaxBlocks * 4 + 255) / 256 * 256;
constexpr size_t kMcScratchTotal = kMcScratchOffs + (size_t)kMcNoAllocMaxBlocks * 8;
constexpr size_t kMcScratchBytes = kMcScratchTotal + 256;   // 12 MiB + 256 B
constexpr int kMcMaxDevices = 64;

__device__ __align__(256) uint8_t mc_table[kMcTableBytes];
__device__ __align__(256) uint8_t scratch[kMcScratchBytes];

// Device addresses of the two arrays per device (from hipGetSymbolAddress;
// host code never passes the arrays' names to a kernel or a copy).
std::mutex g_mc_mu;
uint8_t* g_mc_table_ptr[kMcMaxDevices] = {};
uint8_t* g_mc_scratch_ptr[kMcMaxDevices] = {};

struct McArgs {
    MkVolView v;
    int x0, y0, z0;
    int cx, cy, cz;  // cubes per axis
    uint64_t ncubes;
    float iso;
    const int8_t* tri;    // [256][16]
    const uint8_t* cnt;   // [256]
};

__device__ inline void cube_coords(const McArgs& A, uint64_t q, int& i, int& j, int& k) {
    i = A.x0 + (int)(q % (uint64_t)A.cx);
    uint64_t r = q / (uint64_t)A.cx;
    j = A.y0 + (int)(r % (uint64_t)A.cy);
    k = A.z0 + (int)(r / (uint64_t)A.cy);
}

// The corner values and case index of cube (i, j, k) (valid: the lane holds a
// cube). Lanes hold consecutive cubes, so a cube's four corners at x = i + 1
// are the next lane's corners at x = i: a shuffle brings them unless the next
// lane is outside the wave or starts another row. Every lane of the wave
// calls this.
__device__ inline int cube_case(const McArgs& A, bool valid, int i, int j, int k, float f[8]) {
    float lo[2][2], hi[2][2];   // corners at x = i and x = i + 1, [dz][dy]
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
            lo[dz][dy] = valid ? (float)A.v.vox[mk_voxel_offset(A.v, i, j + dy, k + dz)] : 0.0f;
    const bool from_next = valid && (threadIdx.x & 63) != 63 && i + 1 < A.x0 + A.cx;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const float nxt = __shfl_down(lo[dz][dy], 1);
            hi[dz][dy] = from_next ? nxt : (valid ? (float)A.v.vox[mk_voxel_offset(A.v, i + 1, j + dy, k + dz)] : 0.0f);
        }
    int c = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int dy = kMkMcCorner[n][1], dz = kMkMcCorner[n][2];
        f[n] = kMkMcCorner[n][0] ? hi[dz][dy] : lo[dz][dy];
        c |= (f[n] < A.iso ? 1 : 0) << n;
    }
    return c;
}

// Block-wide exclusive scan of one value per thread (256 threads, 4 waves).
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    return before + x - v;
}

__global__ void __launch_bounds__(kMcBlock) k_mc_count(McArgs A, uint32_t* __restrict__ block_sum) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint8_t s_cnt[256];
    static_assert(kMcBlock == 256, "one thread stages each case's count");
    s_cnt[threadIdx.x] = A.cnt[threadIdx.x];
    const uint64_t q = (uint64_t)blockIdx.x * kMcBlock + threadIdx.x;
    const bool valid = q < A.ncubes;
    int i = 0, j = 0, k = 0;
    if (valid) cube_coords(A, q, i, j, k);
    float f[8];
    const int c = cube_case(A, valid, i, j, k, f);
    __syncthreads();
    const uint32_t n = valid ? s_cnt[c] : 0u;
    uint32_t total;
    (void)block_exclusive_scan(n, s_wave, total);
    if (threadIdx.x == 0) block_sum[blockIdx.x] = total;
}

// Exclusive scan of the block counts by one workgroup of 1024 threads; writes
// the total. Thread t owns a contiguous run of the counts: it sums its run,
// the run sums are scanned across the workgroup, and the thread writes its
// run's offsets from its run's start.
__global__ void __launch_bounds__(1024) k_mc_scan(const uint32_t* __restrict__ block_sum, uint64_t* __restrict__ block_off,
                                                  uint32_t nblocks, unsigned long long* __restrict__ total,
                                                  unsigned long long* __restrict__ total2) {
    __shared__ unsigned long long s_wave[16];
    const uint32_t per = (nblocks + 1023) / 1024;
    const uint32_t lo = min(nblocks, threadIdx.x * per), hi = min(nblocks, lo + per);
    unsigned long long run = 0;
    // unrolled so that a batch of independent loads shares one round trip
#pragma unroll 16
    for (uint32_t b = lo; b < hi; ++b) run += block_sum[b];
    // inclusive scan of the run sums: within each wave, then over the waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long x = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    unsigned long long before = 0;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    unsigned long long acc = before + x - run;
#pragma unroll 16
    for (uint32_t b = lo; b < hi; ++b) {
        block_off[b] = acc;
        acc += block_sum[b];
    }
    if (threadIdx.x == 1023) {
        *total = before + x;
        if (total2) *total2 = before + x;
    }
}

// Edge e of the cube as used by the emit pass: its corner with the smaller
// coordinate along the edge (lo), the other (hi), the edge's axis, and lo's
// offsets, packed as lo | hi << 3 | axis << 6 | dx << 8 | dy << 9 | dz << 10.
__device__ inline uint32_t mc_edge_info(int e) {
    int a = kMkMcEdge[e][0], b = kMkMcEdge[e][1];
    const bool swap = kMkMcCorner[a][0] + kMkMcCorner[a][1] + kMkMcCorner[a][2] >
                      kMkMcCorner[b][0] + kMkMcCorner[b][1] + kMkMcCorner[b][2];
    const int lo = swap ? b : a, hi = swap ? a : b;
    const int axis = kMkMcCorner[lo][0] != kMkMcCorner[hi][0] ? 0 : (kMkMcCorner[lo][1] != kMkMcCorner[hi][1] ? 1 : 2);
    return (uint32_t)lo | (uint32_t)hi << 3 | (uint32_t)axis << 6 | (uint32_t)kMkMcCorner[lo][0] << 8 |
           (uint32_t)kMkMcCorner[lo][1] << 9 | (uint32_t)kMkMcCorner[lo][2] << 10;
}

// f[n] for a lane-varying corner index n, by selects (no indexed registers).
__device__ inline float mc_corner(const float f[8], uint32_t n) {
    const float a = (n & 1) ? f[1] : f[0], b = (n & 1) ? f[3] : f[2];
    const float c = (n & 1) ? f[5] : f[4], d = (n & 1) ? f[7] : f[6];
    const float ab = (n & 2) ? b : a, cd = (n & 2) ? d : c;
    return (n & 4) ? cd : ab;
}

// The triangle table (edge indices per case), the triangle counts and the
// edge table are staged in LDS, so that a vertex costs one LDS byte and one
// LDS word instead of a chain of dependent global loads.
__global__ void __launch_bounds__(kMcBlock) k_mc_emit(McArgs A, const uint64_t* __restrict__ block_off,
                                                      float* __restrict__ out, uint64_t capacity) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint4 s_tri[256];      // [case] 16 edge indices
    __shared__ uint8_t s_cnt[256];
    __shared__ uint32_t s_edge[12];
    static_assert(kMcBlock == 256, "one thread stages each case's row and count");
    s_tri[threadIdx.x] = ((const uint4*)A.tri)[threadIdx.x];
    s_cnt[threadIdx.x] = A.cnt[threadIdx.x];
    if (threadIdx.x < 12) s_edge[threadIdx.x] = mc_edge_info((int)threadIdx.x);
    const uint64_t q = (uint64_t)blockIdx.x * kMcBlock + threadIdx.x;
    const bool valid = q < A.ncubes;
    float f[8];
    int i = 0, j = 0, k = 0;
    if (valid) cube_coords(A, q, i, j, k);
    const int c = cube_case(A, valid, i, j, k, f);
    __syncthreads();
    const uint32_t n = valid ? s_cnt[c] : 0u;
    uint32_t total;
    const uint32_t rank = block_exclusive_scan(n, s_wave, total);
    uint64_t t = block_off[blockIdx.x] + rank;
    const int8_t* row = (const int8_t*)&s_tri[c];
    for (uint32_t m = 0; m < n; ++m, ++t) {
        if (t >= capacity) break;
        float* o = out + 9 * t;
        for (int vtx = 0; vtx < 3; ++vtx) {
            const uint32_t info = s_edge[row[3 * m + vtx]];
            const float flo = mc_corner(f, info & 7u), fhi = mc_corner(f, (info >> 3) & 7u);
            const float tt = (A.iso - flo) / (fhi - flo);
            float p[3] = {(float)(i + (int)((info >> 8) & 1u)), (float)(j + (int)((info >> 9) & 1u)),
                          (float)(k + (int)((info >> 10) & 1u))};
            const uint32_t axis = (info >> 6) & 3u;
            p[0] = axis == 0 ? p[0] + tt : p[0];
            p[1] = axis == 1 ? p[1] + tt : p[1];
            p[2] = axis == 2 ? p[2] + tt : p[2];
            o[3 * vtx + 0] = p[0];
            o[3 * vtx + 1] = p[1];
            o[3 * vtx + 2] = p[2];
        }
    }
}

mk_status prepare(mk_context* ctx, const int* box, float iso, McArgs& A, uint32_t& nblocks) {
    if (!box) return mk_fail(ctx, MK_ERR_ARG, "null box");
    if (!ctx->d_vox) return mk_fail(ctx, MK_ERR_STATE, "marching cubes before any volume was uploaded");
    if (!std::isfinite(iso)) return mk_fail(ctx, MK_ERR_ARG, "iso value is not finite");
    const int x0 = box[0], y0 = box[1], z0 = box[2], x1 = box[3], y1 = box[4], z1 = box[5];
    if (x0 < 0 || y0 < 0 || z0 < 0 || x1 > ctx->nx || y1 > ctx->ny || z1 > ctx->nz || x1 - x0 < 2 || y1 - y0 < 2 ||
        z1 - z0 < 2)
        return mk_fail(ctx, MK_ERR_ARG, "bad box [%d,%d,%d)-[%d,%d,%d) for volume %dx%dx%d", x0, y0, z0, x1, y1, z1,
                       ctx->nx, ctx->ny, ctx->nz);
    uint8_t* table = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mc_mu);
        if (ctx->device >= 0 && ctx->device < kMcMaxDevices) table = g_mc_table_ptr[ctx->device];
    }
    if (!table) return mk_fail(ctx, MK_ERR_STATE, "marching-cubes table not set up on device %d", ctx->device);
    A.v = ctx->view();
    A.x0 = x0; A.y0 = y0; A.z0 = z0;
    A.cx = x1 - x0 - 1; A.cy = y1 - y0 - 1; A.cz = z1 - z0 - 1;
    A.ncubes = (uint64_t)A.cx * A.cy * A.cz;
    A.iso = iso;
    A.tri = (const int8_t*)table;
    A.cnt = table + 256 * 16;
    const uint64_t nb = (A.ncubes + kMcBlock - 1) / kMcBlock;
    if (nb > 0x7FFFFFFFull) return mk_fail(ctx, MK_ERR_ARG, "box too large");
    nblocks = (uint32_t)nb;
    return MK_OK;
}

// The context's scratch: per-block counts, then 64-byte aligned offsets.
uint64_t* scratch_offsets(void* base, uint32_t nblocks) {
    return (uint64_t*)(((uintptr_t)((uint32_t*)base + nblocks) + 63) & ~(uintptr_t)63);
}

// Calls of one context share its scratch: s waits on the device for the last
// call's kernels when that call ran on another stream; the scratch grows here.
mk_status acquire_scratch(mk_context* ctx, uint32_t nblocks, hipStream_t s) {
    if (ctx->mc_used && ctx->mc_stream != s) MK_HIP(ctx, hipStreamWaitEvent(s, ctx->mc_done, 0));
    mk_status st = mk_ensure_buffer(ctx, &ctx->d_mc_scratch, &ctx->mc_scratch_cap,
                                    (size_t)nblocks * (sizeof(uint32_t) + sizeof(uint64_t)) + 64);
    if (st) return st;
    if (!ctx->d_mc_total) MK_HIP(ctx, hipMalloc((void**)&ctx->d_mc_total, 64));
    return MK_OK;
}

mk_status release_scratch(mk_context* ctx, hipStream_t s) {
    MK_HIP(ctx, hipEventRecord(ctx->mc_done, s));
    ctx->mc_stream = s;
    ctx->mc_used = true;
    return MK_OK;
}

}  // namespace

// Sets up, once per device (with that device current), the device addresses
// of the triangle table and of the allocation-free entry point's scratch, and
// fills the table.
mk_status mk_mc_device_init(int device) {
    if (device < 0 || device >= kMcMaxDevices)
        return mk_fail(nullptr, MK_ERR_ARG, "marching cubes: device %d beyond %d", device, kMcMaxDevices);
    std::lock_guard<std::mutex> lk(g_mc_mu);
    if (g_mc_table_ptr[device]) return MK_OK;
    void* table = nullptr;
    void* scr = nullptr;
    hipError_t e = hipGetSymbolAddress(&table, HIP_SYMBOL(mc_table));
    if (e == hipSuccess) e = hipGetSymbolAddress(&scr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !table || !scr)
        return mk_fail(nullptr, MK_ERR_HIP, "hipGetSymbolAddress (marching cubes, device %d): %s", device,
                       hipGetErrorString(e));
    int8_t tri[256][16];
    uint8_t cnt[256];
    mk_mc_expand(tri, cnt);
    uint8_t buf[kMcTableBytes];
    memcpy(buf, tri, 256 * 16);
    memcpy(buf + 256 * 16, cnt, 256);
    e = hipMemcpy(table, buf, sizeof(buf), hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return mk_fail(nullptr, MK_ERR_HIP, "marching-cubes table upload (device %d): %s", device, hipGetErrorString(e));
    g_mc_table_ptr[device] = (uint8_t*)table;
    g_mc_scratch_ptr[device] = (uint8_t*)scr;
    return MK_OK;
}

extern "C" mk_status mk_marching_cubes_async(mk_context* ctx, const int* box, float iso, mk_triangle* dev_out,
                                             uint64_t capacity, uint64_t* dev_count, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    McArgs A;
    uint32_t nblocks = 0;
    mk_status st = prepare(ctx, box, iso, A, nblocks);
    if (st) return st;
    if (capacity > 0 && !dev_out) return mk_fail(ctx, MK_ERR_ARG, "null triangle buffer");
    hipStream_t s = (hipStream_t)stream;
    st = mk_order_after_context(ctx, s);
    if (st) return st;
    st = acquire_scratch(ctx, nblocks, s);
    if (st) return st;
    uint32_t* sums = (uint32_t*)ctx->d_mc_scratch;
    uint64_t* offs = scratch_offsets(ctx->d_mc_scratch, nblocks);
    hipLaunchKernelGGL(k_mc_count, dim3(nblocks), dim3(kMcBlock), 0, s, A, sums);
    hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, s, sums, offs, nblocks,
                       (unsigned long long*)ctx->d_mc_total, (unsigned long long*)dev_count);
    if (capacity > 0)
        hipLaunchKernelGGL(k_mc_emit, dim3(nblocks), dim3(kMcBlock), 0, s, A, offs, (float*)dev_out, capacity);
    MK_HIP(ctx, hipGetLastError());
    st = release_scratch(ctx, s);
    if (st) return st;
    return mk_note_reader(ctx, s);
}

extern "C" mk_status mk_marching_cubes_noalloc(mk_context* ctx, const int* box, float iso, mk_triangle* dev_out,
                                               uint64_t capacity, uint64_t* dev_count, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    McArgs A;
    uint32_t nblocks = 0;
    mk_status st = prepare(ctx, box, iso, A, nblocks);
    if (st) return st;
    if (nblocks > kMcNoAllocMaxBlocks)
        return mk_fail(ctx, MK_ERR_ARG,
                       "mk_marching_cubes_noalloc: box of %llu cubes exceeds the scratch limit of %llu cubes (%u blocks "
                       "of %d)",
                       (unsigned long long)A.ncubes, (unsigned long long)kMcNoAllocMaxBlocks * kMcBlock,
                       kMcNoAllocMaxBlocks, kMcBlock);
    if (capacity > 0 && !dev_out) return mk_fail(ctx, MK_ERR_ARG, "null triangle buffer");
    uint8_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mc_mu);
        base = g_mc_scratch_ptr[ctx->device];
    }
    if (!base) return mk_fail(ctx, MK_ERR_STATE, "marching-cubes scratch not set up on device %d", ctx->device);
    // every block's count, every offset and the total are written before they
    // are read, so nothing needs clearing
    uint32_t* sums = (uint32_t*)base;
    uint64_t* offs = (uint64_t*)(base + kMcScratchOffs);
    unsigned long long* total = (unsigned long long*)(base + kMcScratchTotal);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mc_count, dim3(nblocks), dim3(kMcBlock), 0, s, A, sums);
    hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, s, sums, offs, nblocks, total,
                       (unsigned long long*)dev_count);
    if (capacity > 0)
        hipLaunchKernelGGL(k_mc_emit, dim3(nblocks), dim3(kMcBlock), 0, s, A, offs, (float*)dev_out, capacity);
    MK_HIP(ctx, hipGetLastError());
    return MK_OK;
}

extern "C" mk_status mk_marching_cubes(mk_context* ctx, const int* box, float iso, mk_triangle* host_out,
                                       uint64_t capacity, uint64_t* n_triangles) {
    if (!ctx) return MK_ERR_ARG;
    if (!n_triangles) return mk_fail(ctx, MK_ERR_ARG, "null count");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    McArgs A;
    uint32_t nblocks = 0;
    mk_status st = prepare(ctx, box, iso, A, nblocks);
    if (st) return st;
    st = acquire_scratch(ctx, nblocks, ctx->stream);
    if (st) return st;
    uint32_t* sums = (uint32_t*)ctx->d_mc_scratch;
    uint64_t* offs = scratch_offsets(ctx->d_mc_scratch, nblocks);
    hipLaunchKernelGGL(k_mc_count, dim3(nblocks), dim3(kMcBlock), 0, ctx->stream, A, sums);
    hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, ctx->stream, sums, offs, nblocks,
                       (unsigned long long*)ctx->d_mc_total, (unsigned long long*)nullptr);
    MK_HIP(ctx, hipGetLastError());
    uint64_t total = 0;
    MK_HIP(ctx, hipMemcpyAsync(&total, ctx->d_mc_total, sizeof(total), hipMemcpyDeviceToHost, ctx->stream));
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *n_triangles = total;
    if (!host_out || capacity == 0 || total == 0) return release_scratch(ctx, ctx->stream);
    const uint64_t n = total < capacity ? total : capacity;
    void* d_out = nullptr;
    MK_HIP(ctx, hipMalloc(&d_out, n * sizeof(mk_triangle)));
    hipLaunchKernelGGL(k_mc_emit, dim3(nblocks), dim3(kMcBlock), 0, ctx->stream, A, offs, (float*)d_out, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, d_out, n * sizeof(mk_triangle), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) return mk_hip_fail(ctx, e, "marching cubes emit");
    return release_scratch(ctx, ctx->stream);
}
