// mk_batch_dev.h -- descriptor layout and item lookup shared by the batch
// kernels. The render variants (mk_batch.hip) and the MPR/CPR variant
// (mk_batch_plane.hip) live in separate translation units so that each is
// compiled with the floating-point contraction flags of its one-request twin
// (mk_render.hip, mk_plane.hip): identical flags give identical images.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "mk_plane_dev.h"
#include "mk_render_dev.h"

namespace mkb {

constexpr int kTile = 16;
constexpr int kBatchLabelSlots = 16;   // per-label accumulators per thread in LDS
constexpr int kBatchLutLds = 1024;     // LUT entries staged in LDS

enum : int { F_NEAREST = 1 };

// Kernel variants; an item's variant fixes the template arguments of its
// per-pixel code.
enum Variant : int {
    V_VR_MASK_LDS,
    V_VR_MASK_GLB,
    V_VR_LDS,
    V_VR_GLB,
    V_MIP_MASK,
    V_MIP,
    V_SURFACE,
    V_PLANE_CPR,
    V_COUNT
};

constexpr size_t lds_bytes(int v) {
    return v == V_VR_MASK_LDS || v == V_VR_MASK_GLB
               ? mkr::kLabelLdsBytes + kBatchLutLds * sizeof(float4) + kBatchLabelSlots * 256 * sizeof(float)
           : v == V_VR_LDS || v == V_VR_GLB ? mkr::kLabelLdsBytes + kBatchLutLds * sizeof(float4)
           : v == V_MIP_MASK || v == V_MIP ? mkr::kLabelLdsBytes
                                            : 0;
}

struct BatchDesc {
    int kind;      // MK_ITEM_*
    int sub;       // MPR type for planes
    int flags;
    int tiles_x;
    int width, height;
    float4 aw[68];
    union {
        mkr::RenderArgs ra;
        mkp::PlaneArgs pa;
        mkp::CprArgs ca;
    };
};

struct VariantRange {
    int first, count;   // items [first, first + count) of the sorted descriptor array
    uint32_t tiles;
    uint32_t pad;
};

struct BatchHeader {
    int n_items;
    int pad;
    VariantRange var[V_COUNT];
};

// Finds the item and pixel of this thread for variant V; false when the
// workgroup lies past the variant's tiles.
template <int V>
__device__ inline bool batch_item(const uint8_t* __restrict__ block, uint32_t tile_base_off, uint32_t desc_off,
                                  const BatchDesc*& item, int& x, int& y) {
    const BatchHeader& H = *(const BatchHeader*)block;
    const VariantRange R = H.var[V];
    const uint32_t g = blockIdx.x;
    if (g >= R.tiles) return false;
    const uint32_t* base = (const uint32_t*)(block + tile_base_off);
    const BatchDesc* items = (const BatchDesc*)(block + desc_off);
    int lo = R.first, hi = R.first + R.count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    // lo is the same for the whole workgroup; saying so lets the descriptor
    // fields load through the scalar unit
    lo = __builtin_amdgcn_readfirstlane(lo);
    item = items + lo;
    // the item's tiles in rows per XCD, as a one-request launch maps them
    // (workgroup g runs on XCD g % 8, so equal (g - base) % 8 share an XCD)
    const uint32_t t = mk_xcd_row_tile(g - base[lo], (uint32_t)item->tiles_x,
                                       (uint32_t)((item->height + kTile - 1) / kTile));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    x = (int)(t % (uint32_t)item->tiles_x) * kTile + (wave & 1) * 8 + (lane & 7);
    y = (int)(t / (uint32_t)item->tiles_x) * kTile + (wave >> 1) * 8 + (lane >> 3);
    return true;
}

// Kernel of the MPR/CPR variant (defined in mk_batch_plane.hip).
const void* plane_cpr_kernel();

}  // namespace mkb
