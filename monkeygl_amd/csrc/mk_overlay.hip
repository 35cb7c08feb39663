// mk_overlay.hip -- the orientation figure drawn into VR images.
//
// The reference draws a small human figure, seen through the VR rotation, into
// the bottom-right 64x64 corner of every VR image on the host
// (VRImageLayer::MergeOrientationBox / GetMeshPoints, core/VRImageLayer.cpp:
// 104-194): the figure's facets are rotated, back faces dropped, each facet
// rasterised into a 512x512 float image with a depth buffer
// (Methods::FillHoleInImage_Ch1, core/Methods.cpp:308-368), and every 8th
// pixel of that image lights the corner. The mesh is the binary STL the
// reference loads at start-up (MarchingCube::LoadBinarySTL,
// core/MarchingCube.cpp:306-365).
//
// Here the corner is computed directly at its 64x64 sample points. A facet
// paints the pixels of a row of its bounding box from the first to the last
// pixel of its three edge lines in that row, both included (the row scans of
// FillHoleInImage_Ch1 mark only the zeros outside the edges, and the edges are
// drawn with brightness 255); the edge pixels are the reference's line dots
// (DrawLineInImage8Bit of width 1: steps of half a pixel, positions truncated
// after adding 0.5). The sequential depth test
// (paint when the stored depth is 0 or smaller) leaves each pixel to the first
// facet with the largest depth, which one 64-bit atomic maximum of (depth
// bits, inverted facet index) per sample point reproduces in parallel. A
// thread handles a facet; the few facets spanning many sample rows are
// queued and get a wave each, one lane per row.
//
// The arithmetic follows the reference's types and order (double rotation
// sums, float directions and depth), so the file is built without
// floating-point contraction.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstring>
#include <vector>

#include "mk_context.h"

namespace {

constexpr int kRaster = 512;   // the reference's figure image
constexpr int kBox = 64;       // corner size
constexpr int kStep = 8;       // kRaster / kBox

// One edge line of a facet in the facet's local mask coordinates.
struct Line {
    float x0, y0;
    double dx, dy;   // the normalised direction
    int nsteps;      // dots at i = 0..nsteps; -1 draws nothing
};

__device__ inline Line make_line(float x0, float y0, float x1, float y1) {
    Line L;
    L.x0 = x0;
    L.y0 = y0;
    const float fx = x1 - x0, fy = y1 - y0;
    const double len = (double)sqrtf(fx * fx + fy * fy);
    if (fx == 0.0f && fy == 0.0f) {
        // zero length: the normalised direction is NaN and no dot lands
        L.dx = L.dy = 0.0;
        L.nsteps = -1;
        return L;
    }
    const double ex = (double)fx, ey = (double)fy;
    const double n = sqrt(ex * ex + ey * ey);
    L.dx = ex / n;
    L.dy = ey / n;
    L.nsteps = (int)(len + 0.5) * 2;
    return L;
}

__device__ inline int dot_x(const Line& L, int i) { return (int)((double)L.x0 + i * L.dx / 2.0 + 0.5); }
__device__ inline int dot_y(const Line& L, int i) { return (int)((double)L.y0 + i * L.dy / 2.0 + 0.5); }

// The x extent [lo, hi] of the line's dots in local row yl, clipped to the
// mask width; false when no dot lands there. Both dot coordinates are
// monotone in i and move by at most one pixel per half-pixel step, so a
// row's dots are one run of i, found by bisection, and their x values cover
// every integer between the run's end points.
__device__ inline bool row_extent(const Line& L, int yl, int nw, int& lo, int& hi) {
    if (L.nsteps < 0) return false;
    const int ya = dot_y(L, 0), yb = dot_y(L, L.nsteps);
    const bool up = yb >= ya;
    if (up ? (yl < ya || yl > yb) : (yl > ya || yl < yb)) return false;
    int a = 0, b = L.nsteps;   // first i at or past yl
    while (a < b) {
        const int m = (a + b) >> 1;
        const int y = dot_y(L, m);
        if (up ? y >= yl : y <= yl) b = m;
        else a = m + 1;
    }
    const int i0 = a;
    b = L.nsteps;              // last i at or before yl
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        const int y = dot_y(L, m);
        if (up ? y <= yl : y >= yl) a = m;
        else b = m - 1;
    }
    const int xa = dot_x(L, i0), xb = dot_x(L, a);
    const int xl = max(min(xa, xb), 0), xh = min(max(xa, xb), nw - 1);
    if (xl > xh) return false;
    lo = xl;
    hi = xh;
    return true;
}

__device__ inline float3 rotate(const float* m, float3 p) {
    // Methods::GetTransferPointf: double sums of float products, then float
    double r[3];
    for (int i = 0; i < 3; ++i) {
        r[i] = 0;
        r[i] += (double)m[i * 3 + 0] * (double)p.x;
        r[i] += (double)m[i * 3 + 1] * (double)p.y;
        r[i] += (double)m[i * 3 + 2] * (double)p.z;
    }
    return make_float3((float)r[0], (float)r[1], (float)r[2]);
}

__device__ inline float3 fdir(float x, float y, float z) {
    // Direction3f(x, y, z): divided by its float length
    const float n = sqrtf(x * x + y * y + z * z);
    return make_float3(x / n, y / n, z / n);
}

struct Rot9 {
    float m[9];
};

// Per-call state in the workspace: the sample-point keys, each facet's
// diffuse term, and the queue of facets spanning many sample rows.
struct OrientWs {
    unsigned long long* keys;   // kBox * kBox
    float* diffuse;             // n
    unsigned int* nbig;         // queue length
    int* big;                   // n
};

__global__ void __launch_bounds__(256) k_clear_keys(OrientWs ws) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kBox * kBox) ws.keys[i] = 0ull;
    if (i == 0) *ws.nbig = 0u;
}

// A facet after rotation, culling and projection, ready for its sample rows.
struct FacetRows {
    Line L[3];
    unsigned long long key;
    int xmin, ymin, nw;
    int y0, y1;   // sample rows y0, y0 + 8, ..., <= y1
};

__device__ inline bool facet_setup(const float* __restrict__ facets, int f, const Rot9& rot, FacetRows& R, float& d) {
    const float* q = facets + (size_t)f * 9;
    const float3 v1 = rotate(rot.m, make_float3(q[0], q[1], q[2]));
    const float3 v2 = rotate(rot.m, make_float3(q[3], q[4], q[5]));
    const float3 v3 = rotate(rot.m, make_float3(q[6], q[7], q[8]));
    // n = Direction3f(v3, v1).cross(Direction3f(v3, v2)); d = n . (0, -1, 0)
    const float3 a = fdir(v1.x - v3.x, v1.y - v3.y, v1.z - v3.z);
    const float3 b = fdir(v2.x - v3.x, v2.y - v3.y, v2.z - v3.z);
    const float3 nn = fdir(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
    d = 0.0f;
    d += nn.x * 0.0f;
    d += nn.y * -1.0f;
    d += nn.z * 0.0f;
    if (d < 0.0f) return false;   // NaN passes, as in the reference
    const float zbuf = (v1.y + v2.y + v3.y) / 3 + 10000;
    // TransferPoint2D: (x, z) * 512 + (256, 256)
    const float r = (float)kRaster, c = (float)(kRaster / 2.0);
    const float p1x = (float)((double)v1.x * r) + c, p1y = (float)((double)v1.z * r) + c;
    const float p2x = (float)((double)v2.x * r) + c, p2y = (float)((double)v2.z * r) + c;
    const float p3x = (float)((double)v3.x * r) + c, p3y = (float)((double)v3.z * r) + c;
    // FillHoleInImage_Ch1's bounding box
    const float fxmin = fminf(p1x, fminf(p2x, p3x)), fymin = fminf(p1y, fminf(p2y, p3y));
    const float fxmax = fmaxf(p1x, fmaxf(p2x, p3x)), fymax = fmaxf(p1y, fmaxf(p2y, p3y));
    // a facet beyond +-1e6 pixels (or with non-finite vertices) is skipped:
    // there the reference's integer conversions are undefined and its mask of
    // the bounding box would not fit in memory; it also bounds every loop below
    const float lim = 1.0e6f;
    if (!(fabsf(fxmin) < lim && fabsf(fymin) < lim && fabsf(fxmax) < lim && fabsf(fymax) < lim)) return false;
    const int xmin = (int)fxmin, ymin = (int)fymin;
    const int xmax = (int)((double)fxmax + 0.5), ymax = (int)((double)fymax + 0.5);
    if (xmin >= kRaster || ymin >= kRaster || xmax < 0 || ymax < 0) return false;
    const int nw = xmax - xmin + 1, nh = ymax - ymin + 1;
    if ((long long)nw * nh > (1ll << 26)) return false;
    const float ox = (float)xmin, oy = (float)ymin;
    R.L[0] = make_line(p1x - ox, p1y - oy, p2x - ox, p2y - oy);
    R.L[1] = make_line(p1x - ox, p1y - oy, p3x - ox, p3y - oy);
    R.L[2] = make_line(p2x - ox, p2y - oy, p3x - ox, p3y - oy);
    R.key = ((unsigned long long)__float_as_uint(zbuf) << 32) | (0xFFFFFFFFu - (uint32_t)f);
    R.xmin = xmin;
    R.ymin = ymin;
    R.nw = nw;
    // sample rows Y = 8k inside the box rows and the raster
    int y0 = ymin > 0 ? ymin : 0;
    R.y0 = (y0 + kStep - 1) / kStep * kStep;
    R.y1 = ymax < kRaster - 1 ? ymax : kRaster - 1;
    return true;
}

// One sample row Y of a facet: the painted sample points, lo <= x <= hi in
// the facet's local row, inside the raster.
__device__ inline void facet_row(const FacetRows& R, int Y, unsigned long long* __restrict__ keys) {
    const int yl = Y - R.ymin;
    int lo = INT_MAX, hi = INT_MIN;
    for (int e = 0; e < 3; ++e) {
        int a, b;
        if (row_extent(R.L[e], yl, R.nw, a, b)) {
            lo = min(lo, a);
            hi = max(hi, b);
        }
    }
    if (lo > hi) return;   // no edge pixel in this row
    int X0 = lo + R.xmin;
    X0 = X0 > 0 ? X0 : 0;
    X0 = (X0 + kStep - 1) / kStep * kStep;
    for (int X = X0; X <= hi + R.xmin && X < kRaster; X += kStep)
        atomicMax(&keys[(Y / kStep) * kBox + X / kStep], R.key);
}

constexpr int kRowsPerThread = 2;   // facets with more sample rows go to k_orient_big

// One thread per facet; facets spanning more than kRowsPerThread sample rows
// are queued for a wave each.
__global__ void __launch_bounds__(256) k_orient_facets(const float* __restrict__ facets, int n, Rot9 rot, OrientWs ws) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    FacetRows R;
    float d;
    if (!facet_setup(facets, f, rot, R, d)) return;
    ws.diffuse[f] = d;
    if (R.y1 - R.y0 >= kRowsPerThread * kStep) {
        ws.big[atomicAdd(ws.nbig, 1u)] = f;
        return;
    }
    for (int Y = R.y0; Y <= R.y1; Y += kStep) facet_row(R, Y, ws.keys);
}

// One wave per queued facet, one lane per sample row (at most 64 rows).
__global__ void __launch_bounds__(256) k_orient_big(const float* __restrict__ facets, Rot9 rot, OrientWs ws) {
    const unsigned int nbig = *ws.nbig;
    const int lane = threadIdx.x & 63;
    const unsigned int waves = gridDim.x * (blockDim.x >> 6);
    for (unsigned int k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < nbig; k += waves) {
        FacetRows R;
        float d;
        if (!facet_setup(facets, ws.big[k], rot, R, d)) continue;
        const int Y = R.y0 + lane * kStep;
        if (Y <= R.y1) facet_row(R, Y, ws.keys);
    }
}

// The corner: RGBA(0.902, 0.902, 0.302) * diffuse + 0.05, each step clamped
// to [0, 1], times 255 truncated, where the winning facet's diffuse is > 0.
__global__ void __launch_bounds__(256) k_orient_compose(const unsigned long long* __restrict__ keys,
                                                        const float* __restrict__ diffuse, uint8_t* __restrict__ rgb,
                                                        int width, int height) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kBox * kBox) return;
    const unsigned long long key = keys[i];
    if (key == 0ull) return;
    const float d = diffuse[0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFull)];
    if (!(d > 0.0f)) return;
    const int r = i / kBox, c = i % kBox;
    const float base[3] = {0.902f, 0.902f, 0.302f};
    uint8_t* px = rgb + 3 * ((size_t)(height - kBox + r) * width + (width - kBox + c));
    for (int k = 0; k < 3; ++k) {
        float v = base[k] * d;
        v = v < 0 ? 0 : v;
        v = v > 1 ? 1 : v;
        v = v + 0.05f;
        v = v < 0 ? 0 : v;
        v = v > 1 ? 1 : v;
        px[k] = (uint8_t)(int)(v * 255);
    }
}

// A polyline drawn as Methods::DrawLineInImage24Bit of width 1 per segment
// (core/Methods.cpp:183-200): one workgroup per segment, its threads over the
// segment's dots. Every dot gets the same colour, so overlapping writes agree.
__global__ void __launch_bounds__(64) k_polyline(const float* __restrict__ xy, int n, uint8_t r, uint8_t g, uint8_t b,
                                                 uint8_t* __restrict__ rgb, int width, int height) {
    const int k = blockIdx.x;
    if (k + 1 >= n) return;
    const Line L = make_line(xy[2 * k], xy[2 * k + 1], xy[2 * k + 2], xy[2 * k + 3]);
    for (int i = threadIdx.x; i <= L.nsteps; i += blockDim.x) {
        const int x = dot_x(L, i), y = dot_y(L, i);
        if (x < 0 || x >= width || y < 0 || y >= height) continue;
        uint8_t* px = rgb + 3 * ((size_t)y * width + x);
        px[0] = r;
        px[1] = g;
        px[2] = b;
    }
}

}  // namespace

extern "C" mk_status mk_draw_polyline_async(mk_context* ctx, const float* dev_xy, int n, const float* color3,
                                            uint8_t* dev_rgb, int width, int height, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    if (n < 0 || (n > 0 && (!dev_xy || !color3)) || !dev_rgb || width <= 0 || height <= 0)
        return mk_fail(ctx, MK_ERR_ARG, "mk_draw_polyline_async: bad arguments");
    if (n < 2) return MK_OK;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    // DrawDotInImage24Bit: int(c * 255) in float
    const uint8_t r = (uint8_t)(int)(color3[0] * 255), g = (uint8_t)(int)(color3[1] * 255),
                  b = (uint8_t)(int)(color3[2] * 255);
    hipLaunchKernelGGL(k_polyline, dim3((uint32_t)(n - 1)), dim3(64), 0, (hipStream_t)stream, dev_xy, n, r, g, b,
                       dev_rgb, width, height);
    MK_HIP(ctx, hipGetLastError());
    return MK_OK;
}

extern "C" mk_status mk_set_orientation_mesh(mk_context* ctx, const uint8_t* stl_facets, int n_facets) {
    if (!ctx) return MK_ERR_ARG;
    if (n_facets < 0 || (n_facets > 0 && !stl_facets))
        return mk_fail(ctx, MK_ERR_ARG, "mk_set_orientation_mesh: bad facet list");
    MK_HIP(ctx, hipSetDevice(ctx->device));
    // merges on callers' streams may still read the current figure
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->d_orient) (void)hipFree(ctx->d_orient);
    ctx->d_orient = nullptr;
    ctx->orient_n = 0;
    if (n_facets == 0) return MK_OK;
    // MarchingCube::LoadBinarySTL: bounding box (the maxima start at FLT_MIN,
    // as there), centre, scale by the larger of the x and z extents, negate z
    std::vector<float> v((size_t)n_facets * 9);
    float xmin = FLT_MAX, xmax = FLT_MIN, ymin = FLT_MAX, ymax = FLT_MIN, zmin = FLT_MAX, zmax = FLT_MIN;
    for (int i = 0; i < n_facets; ++i) {
        float f[12];
        memcpy(f, stl_facets + (size_t)i * 50, sizeof(f));   // normal, v1, v2, v3
        for (int k = 0; k < 9; ++k) v[(size_t)i * 9 + k] = f[3 + k];
        xmin = std::min(xmin, std::min(f[3], std::min(f[6], f[9])));
        xmax = std::max(xmax, std::max(f[3], std::max(f[6], f[9])));
        ymin = std::min(ymin, std::min(f[4], std::min(f[7], f[10])));
        ymax = std::max(ymax, std::max(f[4], std::max(f[7], f[10])));
        zmin = std::min(zmin, std::min(f[5], std::min(f[8], f[11])));
        zmax = std::max(zmax, std::max(f[5], std::max(f[8], f[11])));
    }
    const float xlen = xmax - xmin, zlen = zmax - zmin;
    const float maxlen = std::max(xlen, zlen);
    const float cx = (xmin + xmax) / 2, cy = (ymin + ymax) / 2, cz = (zmin + zmax) / 2;
    for (size_t j = 0; j < v.size(); j += 3) {
        v[j] = (float)((double)(v[j] - cx) / (double)maxlen);
        v[j + 1] = (float)((double)(v[j + 1] - cy) / (double)maxlen);
        v[j + 2] = -(float)((double)(v[j + 2] - cz) / (double)maxlen);
    }
    MK_HIP(ctx, hipMalloc((void**)&ctx->d_orient, v.size() * sizeof(float)));
    MK_HIP(ctx, hipMemcpy(ctx->d_orient, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->orient_n = n_facets;
    return MK_OK;
}

extern "C" uint64_t mk_orientation_workspace_bytes(mk_context* ctx) {
    if (!ctx) return 0;
    // keys, diffuse (n floats), queue length, queue (n ints), each part 256-byte aligned
    const uint64_t n = (uint64_t)ctx->orient_n;
    return (uint64_t)kBox * kBox * 8 + (n * 4 + 255) / 256 * 256 + 256 + (n * 4 + 255) / 256 * 256;
}

extern "C" mk_status mk_merge_orientation_box_async(mk_context* ctx, const float* rotate9, uint8_t* dev_rgb, int width,
                                                    int height, void* dev_workspace, uint64_t workspace_bytes,
                                                    void* stream) {
    if (!ctx) return MK_ERR_ARG;
    if (!rotate9 || !dev_rgb) return mk_fail(ctx, MK_ERR_ARG, "mk_merge_orientation_box_async: null argument");
    if (width <= kBox || height <= kBox || ctx->orient_n == 0) return MK_OK;
    const uint64_t need = mk_orientation_workspace_bytes(ctx);
    if (!dev_workspace || workspace_bytes < need || ((uintptr_t)dev_workspace & 7) != 0)
        return mk_fail(ctx, MK_ERR_ARG, "mk_merge_orientation_box_async: workspace needs %llu bytes, 8-byte aligned",
                       (unsigned long long)need);
    MK_HIP(ctx, hipSetDevice(ctx->device));
    Rot9 rot;
    memcpy(rot.m, rotate9, sizeof(rot.m));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n = (uint64_t)ctx->orient_n;
    uint8_t* w = (uint8_t*)dev_workspace;
    OrientWs ws;
    ws.keys = (unsigned long long*)w;
    w += (size_t)kBox * kBox * 8;
    ws.diffuse = (float*)w;
    w += (n * 4 + 255) / 256 * 256;
    ws.nbig = (unsigned int*)w;
    w += 256;
    ws.big = (int*)w;
    hipLaunchKernelGGL(k_clear_keys, dim3(kBox * kBox / 256), dim3(256), 0, s, ws);
    hipLaunchKernelGGL(k_orient_facets, dim3((ctx->orient_n + 255) / 256), dim3(256), 0, s, ctx->d_orient,
                       ctx->orient_n, rot, ws);
    hipLaunchKernelGGL(k_orient_big, dim3(64), dim3(256), 0, s, ctx->d_orient, rot, ws);
    hipLaunchKernelGGL(k_orient_compose, dim3(kBox * kBox / 256), dim3(256), 0, s, ws.keys, ws.diffuse, dev_rgb,
                       width, height);
    MK_HIP(ctx, hipGetLastError());
    return mk_note_reader(ctx, s);
}
