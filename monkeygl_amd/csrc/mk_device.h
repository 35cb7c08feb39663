// mk_device.h -- device-side layout and sampling of the bricked volume.
//
// HBM layout. The int16 volume (x fastest on the host, index z*nx*ny+y*nx+x)
// is stored as 8x8x8 bricks of 1 KiB: bricks are ordered x, then y, then z,
// and voxels inside a brick x, then y, then z. One brick is eight 128-byte
// cache lines, so a trilinear footprint (2x2x2) touches at most 8 bricks and a
// wave's 8x8 ray tile stepping through the volume reuses the same lines for
// several steps whatever the ray direction. The uint8 label mask uses the same
// brick order (512 B per brick). Offsets are 32-bit: a volume may hold up to
// 2^31 padded voxels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MK_BRICK_SHIFT 3
#define MK_BRICK 8
#define MK_BRICK_VOXELS 512

struct MkVolView {
    const int16_t* vox;   // bricked int16 voxels
    const uint8_t* mask;  // bricked uint8 labels, or null
    int nx, ny, nz;       // voxel dims
    int nbx, nby, nbz;    // brick dims
};

// Index products are 24-bit multiplies, one full-rate instruction each (a
// 32-bit multiply is quarter-rate). They are exact: a volume holds at most
// 2^31 padded voxels, so brick counts and every partial brick index stay below
// 2^22, and no dimension exceeds 2^17 voxels (MK_MAX_DIM), so a row or slab of
// bricks scaled by 1024 bytes stays below 2^24.
__device__ inline uint32_t mk_brick_index(const MkVolView& v, int x, int y, int z) {
    return __umul24(__umul24((uint32_t)(z >> MK_BRICK_SHIFT), (uint32_t)v.nby) + (uint32_t)(y >> MK_BRICK_SHIFT),
                    (uint32_t)v.nbx) +
           (uint32_t)(x >> MK_BRICK_SHIFT);
}

// Voxel order inside a volume brick: its 2x2x2 blocks of 4x4x4 voxels (128
// bytes, one cache line each) in x, y, z order, x fastest inside a block. The
// 64 samples of a wave lie on a small patch of a plane at any tilt; blocks
// give that patch about half the cache lines that 8x8 z-slices of the brick
// would. Bits of (x, y, z) & 7 in the in-brick offset: x0 x1 y0 y1 z0 z1 x2 y2 z2.
__device__ inline uint32_t mk_inbrick(int x, int y, int z) {
    return ((uint32_t)(z & 4) << 6) | ((uint32_t)(y & 4) << 5) | ((uint32_t)(x & 4) << 4) | ((uint32_t)(z & 3) << 4) |
           ((uint32_t)(y & 3) << 2) | (uint32_t)(x & 3);
}

__device__ inline uint32_t mk_voxel_offset(const MkVolView& v, int x, int y, int z) {
    return (mk_brick_index(v, x, y, z) << 9) | mk_inbrick(x, y, z);
}

// The label mask keeps x-rows inside its bricks (8 bytes per row, z slowest):
// its label kernels handle a brick as 64 rows.
__device__ inline uint32_t mk_mask_offset(const MkVolView& v, int x, int y, int z) {
    return (mk_brick_index(v, x, y, z) << 9) | ((uint32_t)(z & 7) << 6) | ((uint32_t)(y & 7) << 3) | (uint32_t)(x & 7);
}

// The byte offset of a voxel (mk_voxel_offset * 2) as a sum of one term per
// axis, for samplers that combine several indices per axis. Byte offsets fit
// 32 bits (a volume holds at most 2^31 padded voxels), so loads take the
// volume's base address from scalar registers and a 32-bit offset per lane.
__device__ inline uint32_t mk_term_x(int x) {
    return ((uint32_t)(x >> 3) << 10) | ((uint32_t)(x & 4) << 5) | ((uint32_t)(x & 3) << 1);
}
__device__ inline uint32_t mk_term_y(const MkVolView& v, int y) {
    return __umul24((uint32_t)(y >> 3) << 10, (uint32_t)v.nbx) + (((uint32_t)(y & 4) << 6) | ((uint32_t)(y & 3) << 3));
}
__device__ inline uint32_t mk_term_z(const MkVolView& v, int z) {
    return __umul24((uint32_t)(z >> 3) << 10, __umul24((uint32_t)v.nbx, (uint32_t)v.nby)) +
           (((uint32_t)(z & 4) << 7) | ((uint32_t)(z & 3) << 5));
}

// The voxel at byte offset boff (a sum of mk_term_* values).
__device__ inline int mk_vox_at(const MkVolView& v, uint32_t boff) {
    return *(const int16_t*)((const uint8_t*)v.vox + boff);
}

// lo <= hi: max then min (two instructions, where the nested conditional
// compiles to three)
__device__ inline int mk_clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// NaN-safe saturate: NaN maps to 0.
__device__ inline float mk_sat(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// Raw voxel as float with the normalized-float clamp of int16 (-32768 reads
// as -32767): multiplying by 32768/32767 afterwards gives the reference's
// 32768 * tex3D<float>() value.
__device__ inline float mk_raw(const MkVolView& v, int x, int y, int z) {
    int s = v.vox[mk_voxel_offset(v, x, y, z)];
    return (float)(s < -32767 ? -32767 : s);
}

// Split of one normalized coordinate into texel indices and weight:
// x_t = f*N - 0.5, i = floor(x_t), a = x_t - i, neighbours clamped.
__device__ inline void mk_split(float f, int n, int& i0, int& i1, float& a) {
    float xt = f * (float)n - 0.5f;
    float fl = floorf(xt);
    a = xt - fl;
    fl = fminf(fmaxf(fl, -2.0f), (float)(n + 1));
    int i = (int)fl;
    i0 = mk_clampi(i, 0, n - 1);
    i1 = mk_clampi(i + 1, 0, n - 1);
}

__device__ inline float mk_rawo(const MkVolView& v, uint32_t boff) {
    int s = mk_vox_at(v, boff);
    return (float)(s < -32767 ? -32767 : s);
}

// Trilinear sample in HU' units (32768 * normalized-float read). The eight
// corner offsets are sums of two terms per axis (mk_term_*); an index that
// clamps at the volume edge repeats its neighbour's term.
__device__ inline float mk_sample_hu(const MkVolView& v, float fx, float fy, float fz) {
    int x0, x1, y0, y1, z0, z1;
    float ax, ay, az;
    mk_split(fx, v.nx, x0, x1, ax);
    mk_split(fy, v.ny, y0, y1, ay);
    mk_split(fz, v.nz, z0, z1, az);
    const uint32_t X0 = mk_term_x(x0), X1 = mk_term_x(x1);
    const uint32_t Y0 = mk_term_y(v, y0), Y1 = mk_term_y(v, y1);
    const uint32_t Z0 = mk_term_z(v, z0), Z1 = mk_term_z(v, z1);
    float c000 = mk_rawo(v, X0 + Y0 + Z0), c100 = mk_rawo(v, X1 + Y0 + Z0);
    float c010 = mk_rawo(v, X0 + Y1 + Z0), c110 = mk_rawo(v, X1 + Y1 + Z0);
    float c001 = mk_rawo(v, X0 + Y0 + Z1), c101 = mk_rawo(v, X1 + Y0 + Z1);
    float c011 = mk_rawo(v, X0 + Y1 + Z1), c111 = mk_rawo(v, X1 + Y1 + Z1);
    float c00 = c000 + ax * (c100 - c000);
    float c10 = c010 + ax * (c110 - c010);
    float c01 = c001 + ax * (c101 - c001);
    float c11 = c011 + ax * (c111 - c011);
    float c0 = c00 + ay * (c10 - c00);
    float c1 = c01 + ay * (c11 - c01);
    float c = c0 + az * (c1 - c0);
    return c * (32768.0f / 32767.0f);
}

// Point-sampled HU' for nearest MPR/CPR: voxel floor(f*N) clamped, then the
// normalized-float read v/32767 times 32768, in that op order (bit-exact with
// the host restatement).
__device__ inline float mk_sample_nearest_hu(const MkVolView& v, float fx, float fy, float fz) {
    float tx = floorf(fx * (float)v.nx), ty = floorf(fy * (float)v.ny), tz = floorf(fz * (float)v.nz);
    int x = mk_clampi((int)fminf(fmaxf(tx, -1.0f), (float)v.nx), 0, v.nx - 1);
    int y = mk_clampi((int)fminf(fmaxf(ty, -1.0f), (float)v.ny), 0, v.ny - 1);
    int z = mk_clampi((int)fminf(fmaxf(tz, -1.0f), (float)v.nz), 0, v.nz - 1);
    int s = v.vox[mk_voxel_offset(v, x, y, z)];
    float r = (float)s / 32767.0f;
    r = r < -1.0f ? -1.0f : r;
    return 32768.0f * r;
}

// Maps a linear workgroup id to a tile id so that consecutive tiles land on
// the same XCD (workgroups are dispatched round-robin over the 8 XCDs, each
// with its own L2). Bijective on [0, n).
__device__ inline uint32_t mk_xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t X = 8;
    uint32_t q = n / X, r = n % X;
    uint32_t x = b % X, i = b / X;
    return x * q + (x < r ? x : r) + i;
}

// Maps a linear workgroup id to a tile of a tiles_x x tiles_y grid so that XCD
// k (workgroups go round-robin over the 8 XCDs) takes every 8th row of tiles,
// rows k, k + 8, ..., each left to right. A frame's work varies mostly from
// band to band of rows (an object's outline against the background): with
// contiguous bands per XCD the busiest XCD of a body view held 3.5x the work
// of the idlest and set the frame time. Each XCD's rows still cover disjoint
// slabs of the volume. Bijective on [0, tiles_x * tiles_y).
__device__ inline uint32_t mk_xcd_row_tile(uint32_t b, uint32_t tiles_x, uint32_t tiles_y) {
    const uint32_t j = mk_xcd_tile(b, tiles_x * tiles_y);
    // j is a position in the order of rows grouped by their residue mod 8
    uint32_t k = 0, start = 0;
    for (; k < 7; ++k) {
        const uint32_t len = (tiles_y + 7 - k) / 8 * tiles_x;   // rows r < tiles_y with r % 8 == k
        if (j < start + len) break;
        start += len;
    }
    const uint32_t local = j - start;
    return (k + 8 * (local / tiles_x)) * tiles_x + local % tiles_x;
}
