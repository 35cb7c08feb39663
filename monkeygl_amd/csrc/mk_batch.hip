// mk_batch.hip -- batched launches for the render server.
//
// A batch holds up to max_items requests (VR / MIP / Surface frames, slab MPR
// images, CPR images) over any of the batcher's resident volumes, each with
// its own viewport. Items are grouped by kernel variant (VR with/without mask
// and with/without an LDS-staged LUT, MIP with/without mask, Surface, and
// MPR/CPR); one kernel per variant covers the 16x16 tiles of that variant's
// items, a workgroup finding its item by a binary search over the items' first
// tile indices. Every kernel runs the same per-pixel device code as the
// one-request kernels (mk_render_dev.h, mk_plane_dev.h) with the same template
// arguments, so a request renders the same alone, in a batch, and through
// mk_render / mk_render_plane / mk_render_cpr. One kernel per variant, rather
// than one kernel branching over variants, keeps each at the register count
// of its one-request twin (a kernel holding all variants needs ~250 VGPRs).
// A batch of one render request skips the graph and takes the one-request
// launch path (VR in two passes, address-term tables in LDS): at the served
// mix's arrival rate most batches hold one request, and that path renders
// them faster with the same pixels.
//
// Launch: each slot owns a pinned host descriptor block, its device copy and
// a hipGraph built once at creation: [copy the descriptor block] -> [variant 0
// kernel] -> ... -> [variant N-1 kernel]. Submitting fills the pinned block,
// sets the copy size and each present variant's grid (its tile count) in the
// instantiated graph, disables the nodes of absent variants (a disabled node
// passes its dependency through and launches nothing), and replays the graph
// on the caller's stream. A slot is reused only after its
// completion event has fired, so the descriptors of a batch in flight are
// never overwritten. Several slots in flight on separate streams overlap on
// the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "mk_batch_dev.h"
#include "mk_context.h"

namespace {

using namespace mkb;

// Layout of a slot's descriptor block: header, first-tile table (per item,
// relative to its variant's first item), descriptors sorted by variant.
struct BatchLayout {
    size_t tile_base_off, desc_off;
    size_t bytes(int n) const { return desc_off + sizeof(BatchDesc) * (size_t)n; }
};

BatchLayout layout(int max_items) {
    BatchLayout L;
    L.tile_base_off = (sizeof(BatchHeader) + 63) & ~(size_t)63;
    L.desc_off = (L.tile_base_off + sizeof(uint32_t) * (size_t)max_items + 255) & ~(size_t)255;
    return L;
}

template <int V>
__global__ void __launch_bounds__(256) k_batch(const uint8_t* __restrict__ block, uint32_t tile_base_off,
                                               uint32_t desc_off) {
    const BatchDesc* d;
    int x, y;
    if (!batch_item<V>(block, tile_base_off, desc_off, d, x, y)) return;
    if constexpr (V == V_SURFACE) {
        if (x >= d->width || y >= d->height) return;
        mkr::surface_pixel<false>(d->ra, x, y);
    } else {
        constexpr bool kMask = V == V_VR_MASK_LDS || V == V_VR_MASK_GLB || V == V_MIP_MASK;
        constexpr bool kLds = V == V_VR_MASK_LDS || V == V_VR_LDS;
        extern __shared__ float4 smem[];
        mkr::Smem s;
        s.aw = smem;
        s.lab = (int4*)(smem + mkr::kLabelRows);
        s.lut = smem + 2 * mkr::kLabelRows;
        s.aobj = (float*)(s.lut + kBatchLutLds);
        for (int i = threadIdx.x; i < mkr::kLabelRows; i += blockDim.x) s.aw[i] = d->aw[i];
        mkr::stage_labels(d->ra, s.lab);
        if constexpr (kLds)
            for (int i = threadIdx.x; i < d->ra.nlut_lds; i += blockDim.x) s.lut[i] = d->ra.lut[i];
        __syncthreads();
        if (x >= d->width || y >= d->height) return;
        if constexpr (V == V_MIP_MASK || V == V_MIP) mkr::mip_pixel<kMask, false>(d->ra, s, x, y);
        else mkr::vr_pixel<256, kMask, kLds, false>(d->ra, s, x, y, threadIdx.x);
    }
}

const void* variant_kernel(int v) {
    switch (v) {
    case V_VR_MASK_LDS: return (const void*)k_batch<V_VR_MASK_LDS>;
    case V_VR_MASK_GLB: return (const void*)k_batch<V_VR_MASK_GLB>;
    case V_VR_LDS: return (const void*)k_batch<V_VR_LDS>;
    case V_VR_GLB: return (const void*)k_batch<V_VR_GLB>;
    case V_MIP_MASK: return (const void*)k_batch<V_MIP_MASK>;
    case V_MIP: return (const void*)k_batch<V_MIP>;
    case V_SURFACE: return (const void*)k_batch<V_SURFACE>;
    default: return plane_cpr_kernel();
    }
}

}  // namespace

struct mk_batcher {
    std::vector<mk_context*> ctxs;
    int device = 0;
    int max_items = 0;
    uint32_t max_tiles = 0;
    BatchLayout lay{};
    struct Slot {
        uint8_t* h = nullptr;   // pinned descriptor block
        uint8_t* d = nullptr;   // device descriptor block
        uint32_t tile_base_off = 0, desc_off = 0;
        void* args[3] = {};     // kernel arguments (the same for every variant)
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipGraphNode_t copy = nullptr;
        hipGraphNode_t kern[V_COUNT] = {};
        bool enabled[V_COUNT];   // kernel node state in the instantiated graph
        hipEvent_t done = nullptr;
        bool busy = false;
    };
    std::vector<Slot> slots;
    std::vector<int> order;     // submit scratch: item indices sorted by variant
    std::vector<int> variant;
    std::string err = "no error";
};

static mk_status bfail(mk_batcher* b, mk_status st, const std::string& m) {
    b->err = m;
    return st;
}

static mk_status bhip(mk_batcher* b, hipError_t e, const char* what) {
    return bfail(b, e == hipErrorOutOfMemory ? MK_ERR_NOMEM : MK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define B_HIP(b, call)                                      \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return bhip((b), _e, #call);  \
    } while (0)

extern "C" void mk_batcher_destroy(mk_batcher* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (auto& s : b->slots) {
        if (s.done) (void)hipEventSynchronize(s.done);
        if (s.exec) (void)hipGraphExecDestroy(s.exec);
        if (s.graph) (void)hipGraphDestroy(s.graph);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.d) (void)hipFree(s.d);
        if (s.h) (void)hipHostFree(s.h);
    }
    delete b;
}

static hipKernelNodeParams kernel_params(mk_batcher::Slot& s, int v, uint32_t grid) {
    hipKernelNodeParams kp{};
    kp.func = const_cast<void*>(variant_kernel(v));
    kp.gridDim = dim3(grid > 0 ? grid : 1);
    kp.blockDim = dim3(256);
    kp.sharedMemBytes = (unsigned)lds_bytes(v);
    kp.kernelParams = s.args;
    kp.extra = nullptr;
    return kp;
}

static mk_status build_slot(mk_batcher* b, mk_batcher::Slot& s) {
    const size_t full = b->lay.bytes(b->max_items);
    B_HIP(b, hipHostMalloc((void**)&s.h, full, hipHostMallocDefault));
    memset(s.h, 0, full);   // no variant has work until the first submit
    B_HIP(b, hipMalloc((void**)&s.d, full));
    B_HIP(b, hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    s.tile_base_off = (uint32_t)b->lay.tile_base_off;
    s.desc_off = (uint32_t)b->lay.desc_off;
    s.args[0] = &s.d;
    s.args[1] = &s.tile_base_off;
    s.args[2] = &s.desc_off;
    B_HIP(b, hipGraphCreate(&s.graph, 0));
    B_HIP(b, hipGraphAddMemcpyNode1D(&s.copy, s.graph, nullptr, 0, s.d, s.h, full, hipMemcpyHostToDevice));
    hipGraphNode_t prev = s.copy;
    for (int v = 0; v < V_COUNT; ++v) {
        hipKernelNodeParams kp = kernel_params(s, v, 1);
        B_HIP(b, hipGraphAddKernelNode(&s.kern[v], s.graph, &prev, 1, &kp));
        prev = s.kern[v];
    }
    B_HIP(b, hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
    for (int v = 0; v < V_COUNT; ++v) s.enabled[v] = true;
    return MK_OK;
}

static std::string g_batch_create_err = "no error";

extern "C" mk_status mk_batcher_create(mk_context* const* ctxs, int nctx, int max_items, int max_tiles, int nslots,
                                       mk_batcher** out) {
    if (!out) return MK_ERR_ARG;
    *out = nullptr;
    if (!ctxs || nctx <= 0 || max_items <= 0 || max_items > 4096 || max_tiles <= 0 || nslots <= 0 || nslots > 16) {
        g_batch_create_err = "mk_batcher_create: bad arguments";
        return MK_ERR_ARG;
    }
    mk_batcher* b = new mk_batcher();
    b->device = ctxs[0] ? ctxs[0]->device : 0;
    for (int i = 0; i < nctx; ++i) {
        if (!ctxs[i] || ctxs[i]->device != b->device) {
            delete b;
            g_batch_create_err = "mk_batcher_create: null context or contexts on different devices";
            return MK_ERR_ARG;
        }
        b->ctxs.push_back(ctxs[i]);
    }
    b->max_items = max_items;
    b->max_tiles = (uint32_t)max_tiles;
    b->lay = layout(max_items);
    b->slots.resize(nslots);
    b->order.resize(max_items);
    b->variant.resize(max_items);
    mk_status st = MK_OK;
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) st = bhip(b, e, "hipSetDevice");
    for (auto& s : b->slots) {
        if (st) break;
        st = build_slot(b, s);
    }
    if (st) {
        g_batch_create_err = b->err;
        mk_batcher_destroy(b);
        return st;
    }
    *out = b;
    return MK_OK;
}

extern "C" const char* mk_batcher_last_error(const mk_batcher* b) {
    return b ? b->err.c_str() : g_batch_create_err.c_str();
}

extern "C" int mk_batch_query(mk_batcher* b, int slot) {
    if (!b || slot < 0 || slot >= (int)b->slots.size()) return -1;
    auto& s = b->slots[slot];
    if (!s.busy) return 1;
    hipError_t e = hipEventQuery(s.done);
    if (e == hipSuccess) {
        s.busy = false;
        return 1;
    }
    if (e == hipErrorNotReady) return 0;
    bhip(b, e, "hipEventQuery");
    return -1;
}

extern "C" mk_status mk_batch_wait(mk_batcher* b, int slot) {
    if (!b || slot < 0 || slot >= (int)b->slots.size()) return MK_ERR_ARG;
    auto& s = b->slots[slot];
    if (!s.busy) return MK_OK;
    B_HIP(b, hipEventSynchronize(s.done));
    s.busy = false;
    return MK_OK;
}

static mk_status wait_all(mk_batcher* b) {
    for (int i = 0; i < (int)b->slots.size(); ++i) {
        mk_status st = mk_batch_wait(b, i);
        if (st) return st;
    }
    return MK_OK;
}

static bool tables_current(mk_context* c, int type, const mk_alpha_wwwl* aw) {
    const MkSkipTables& sk = c->skip[type];
    return !c->lut_dirty && sk.valid && sk.lut_version == c->lut_version && sk.vol_version == c->vol_version &&
           memcmp(sk.aw, aw, sizeof(sk.aw)) == 0;
}

extern "C" mk_status mk_batch_submit(mk_batcher* b, int slot, const mk_batch_item* items, int n, uint8_t* dev_arena,
                                     uint64_t arena_bytes, void* stream) {
    if (!b) return MK_ERR_ARG;
    if (slot < 0 || slot >= (int)b->slots.size()) return bfail(b, MK_ERR_ARG, "bad slot");
    if (!items || n <= 0 || n > b->max_items) return bfail(b, MK_ERR_ARG, "bad item count " + std::to_string(n));
    if (!dev_arena) return bfail(b, MK_ERR_ARG, "null arena");
    auto& s = b->slots[slot];
    if (s.busy && mk_batch_query(b, slot) != 1) return bfail(b, MK_ERR_STATE, "slot still in flight");
    B_HIP(b, hipSetDevice(b->device));

    // Shared per-context state a batch reads (LUTs, skip tables, step
    // lattice) is rebuilt only after every batch in flight has finished. Items
    // of one batch on one context must agree on that state: one step_scale per
    // context, one alpha/ww/wl table per (context, render type).
    bool rebuild = false;
    const int nctx = (int)b->ctxs.size();
    std::vector<int> first_item(nctx * 4, -1);   // [ctx][0] step_scale, [ctx][1 + type] alpha table
    for (int i = 0; i < n; ++i) {
        const mk_batch_item& it = items[i];
        if (it.context < 0 || it.context >= nctx)
            return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": bad context index");
        mk_context* c = b->ctxs[it.context];
        if (!c->d_vox) return bfail(b, MK_ERR_STATE, "item " + std::to_string(i) + ": context has no volume");
        // the one-request entry points' checks, so an item fails alone and in a batch alike
        mk_status vst = MK_OK;
        if (it.kind == MK_ITEM_RENDER) vst = mk_validate_render(c, &it.render);
        else if (it.kind == MK_ITEM_PLANE) vst = mk_validate_plane(c, &it.plane);
        else if (it.kind == MK_ITEM_CPR) vst = mk_validate_cpr(c, it.cpr_width, it.cpr_height, it.cpr_sampling);
        else return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": unknown kind");
        if (vst) return bfail(b, vst, "item " + std::to_string(i) + ": " + c->err);
        if (it.kind == MK_ITEM_RENDER) {
            const mk_render_params& p = it.render;
            int& fs = first_item[it.context * 4];
            int& fa = first_item[it.context * 4 + 1 + p.render_type];
            if (fs >= 0 && items[fs].render.step_scale != p.step_scale)
                return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": step_scale differs from item " +
                                                std::to_string(fs) + " on the same context");
            if (fa >= 0 && memcmp(items[fa].render.alpha_wwwl, p.alpha_wwwl, sizeof(p.alpha_wwwl)) != 0)
                return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": label alpha/window differs from item " +
                                                std::to_string(fa) + " of the same render type on the same context");
            if (fs < 0) fs = i;
            if (fa < 0) fa = i;
            const float l1 = 1.0f / (float)c->nz * p.step_scale;
            if (!tables_current(c, p.render_type, p.alpha_wwwl) || c->steps_l1 != l1 || !c->d_steps) rebuild = true;
        }
    }
    if (rebuild) {
        mk_status st = wait_all(b);
        if (st) return st;
    }
    for (int i = 0; i < n; ++i) {
        if (items[i].kind != MK_ITEM_RENDER) continue;
        mk_context* c = b->ctxs[items[i].context];
        const mk_render_params& p = items[i].render;
        mk_status st = mk_upload_luts(c);
        if (!st) st = mk_ensure_skip_tables(c, p.render_type, p.alpha_wwwl);
        if (!st) st = mk_ensure_steps(c, 1.0f / (float)c->nz * p.step_scale);
        if (st) return bfail(b, st, std::string("item ") + std::to_string(i) + ": " + c->err);
    }

    // variant of each item, then items grouped by variant (stable)
    for (int i = 0; i < n; ++i) {
        const mk_batch_item& it = items[i];
        int v;
        if (it.kind == MK_ITEM_RENDER) {
            mk_context* c = b->ctxs[it.context];
            const int rt = it.render.render_type;
            const bool mask = c->d_mask != nullptr && rt != MK_RENDER_SURFACE;
            int total_lut = 0;
            for (int l = 0; l < MK_NUM_LABELS; ++l) total_lut = std::max(total_lut, c->lut_tab.off[l] + c->lut_tab.len[l]);
            const bool lds = total_lut <= kBatchLutLds;
            if (rt == MK_RENDER_VR) v = mask ? (lds ? V_VR_MASK_LDS : V_VR_MASK_GLB) : (lds ? V_VR_LDS : V_VR_GLB);
            else if (rt == MK_RENDER_MIP) v = mask ? V_MIP_MASK : V_MIP;
            else v = V_SURFACE;
        } else if (it.kind == MK_ITEM_PLANE || it.kind == MK_ITEM_CPR) {
            v = V_PLANE_CPR;
        } else {
            return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": unknown kind");
        }
        b->variant[i] = v;
        b->order[i] = i;
    }
    std::stable_sort(b->order.begin(), b->order.begin() + n,
                     [&](int x, int y) { return b->variant[x] < b->variant[y]; });

    BatchHeader* H = (BatchHeader*)s.h;
    memset(H, 0, sizeof(*H));
    uint32_t* base = (uint32_t*)(s.h + b->lay.tile_base_off);
    BatchDesc* D = (BatchDesc*)(s.h + b->lay.desc_off);
    uint32_t tiles = 0;
    for (int k = 0; k < n; ++k) {
        const int i = b->order[k];
        const int var = b->variant[i];
        const mk_batch_item& it = items[i];
        mk_context* c = b->ctxs[it.context];
        BatchDesc& d = D[k];
        memset(&d, 0, sizeof(d));
        d.kind = it.kind;
        int w = 0, h = 0;
        size_t out_bytes = 0;
        if (it.kind == MK_ITEM_RENDER) {
            const mk_render_params& p = it.render;
            w = p.width;
            h = p.height;
            out_bytes = (size_t)w * h * 3;
            if (w <= 0 || h <= 0 || w > 16384 || h > 16384)
                return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": bad render size");
            mkr::RenderArgs& A = d.ra;
            A.v = c->view();
            A.lut = c->d_lut;
            A.tab = c->lut_tab;
            const MkSkipTables& sk = c->skip[p.render_type];
            A.occ = sk.occ;
            A.mipb = sk.mipb;
            A.occ2 = sk.occ2;
            A.mipb2 = sk.mipb2;
            A.blabel = c->stats.blabel;
            A.bskip = sk.bskip;
            A.nmx = (c->nbx + 3) / 4;
            A.nmy = (c->nby + 3) / 4;
            A.steps = c->d_steps;
            A.nsteps = c->steps_n;
            A.out = dev_arena + it.out_offset;
            A.width = w;
            A.height = h;
            A.invert_z = p.invert_z != 0;
            for (int q = 0; q < 3; ++q) {
                A.mls[q] = p.max_len_spacing[q];
                A.sv[q] = p.spacing_voxel[q];
            }
            A.xT = p.x_translate;
            A.yT = p.y_translate;
            A.scale = (double)p.scale;
            memcpy(A.M, p.transform, sizeof(A.M));
            A.voi = p.voi;
            memcpy(A.bg, p.background, sizeof(A.bg));
            A.l1 = c->steps_l1;
            A.rl1 = 1.0f / c->steps_l1;
            A.exit_alpha = p.exit_alpha;
            A.max_iter = 24 * c->steps_n + 64;
            const bool mask = var == V_VR_MASK_LDS || var == V_VR_MASK_GLB || var == V_MIP_MASK;
            int nslot = 0;
            for (int l = 0; l < MK_NUM_LABELS; ++l) {
                const bool used = l < 64 ? ((c->used_lo >> l) & 1ull) != 0 : c->used_hi;
                A.slot_of[l] = used ? (uint8_t)nslot : 0;
                if (used) ++nslot;
            }
            A.nslot = mask ? nslot : 0;
            if ((var == V_VR_MASK_LDS || var == V_VR_MASK_GLB) && A.nslot > kBatchLabelSlots)
                return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": more labels than the batch kernel holds");
            int total_lut = 0;
            for (int l = 0; l < MK_NUM_LABELS; ++l) total_lut = std::max(total_lut, c->lut_tab.off[l] + c->lut_tab.len[l]);
            A.nlut_lds = (var == V_VR_MASK_LDS || var == V_VR_LDS) ? total_lut : 0;
            for (int l = 0; l < MK_NUM_LABELS; ++l)
                d.aw[l] = make_float4(p.alpha_wwwl[l].alpha, p.alpha_wwwl[l].ww, p.alpha_wwwl[l].wl, 0.f);
            d.sub = p.render_type;
        } else if (it.kind == MK_ITEM_PLANE) {
            const mk_plane_params& p = it.plane;
            w = p.width;
            h = p.height;
            out_bytes = (size_t)w * h * 2;
            mkp::PlaneArgs& A = d.pa;
            A.v = c->view();
            A.out = (int16_t*)(dev_arena + it.out_offset);
            A.width = w;
            A.height = h;
            A.invert_z = p.invert_z != 0;
            for (int q = 0; q < 3; ++q) {
                A.sp[q] = p.spacing[q];
                A.dh[q] = p.dir_h[q];
                A.dv[q] = p.dir_v[q];
                A.dn[q] = p.dir_n[q];
                A.lt[q] = p.pt_left_top[q];
            }
            A.ps = p.pixel_spacing;
            A.half = p.half_num;
            d.sub = p.mpr_type;
            d.flags = p.sampling == MK_SAMPLE_NEAREST ? F_NEAREST : 0;
        } else {
            w = it.cpr_width;
            h = it.cpr_height;
            out_bytes = (size_t)w * h * 2;
            const uint64_t geo = (uint64_t)h * 3 * sizeof(double);
            if (it.cpr_points_offset % 8 || it.cpr_dirs_offset % 8 || it.cpr_points_offset > arena_bytes ||
                geo > arena_bytes - it.cpr_points_offset || it.cpr_dirs_offset > arena_bytes ||
                geo > arena_bytes - it.cpr_dirs_offset)
                return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": CPR geometry outside the arena");
            mkp::CprArgs& A = d.ca;
            A.v = c->view();
            A.pts = (const double*)(dev_arena + it.cpr_points_offset);
            A.dirs = (const double*)(dev_arena + it.cpr_dirs_offset);
            A.out = (int16_t*)(dev_arena + it.out_offset);
            A.width = w;
            A.height = h;
            A.invert_z = it.cpr_invert_z != 0;
            d.flags = it.cpr_sampling == MK_SAMPLE_NEAREST ? F_NEAREST : 0;
        }
        if (it.out_offset > arena_bytes || out_bytes > arena_bytes - it.out_offset ||
            (it.kind != MK_ITEM_RENDER && it.out_offset % 2))
            return bfail(b, MK_ERR_ARG, "item " + std::to_string(i) + ": output outside the arena");
        d.width = w;
        d.height = h;
        d.tiles_x = (w + kTile - 1) / kTile;
        VariantRange& R = H->var[var];
        if (R.count == 0) R.first = k;
        ++R.count;
        base[k] = R.tiles;
        const uint32_t t = (uint32_t)d.tiles_x * (uint32_t)((h + kTile - 1) / kTile);
        R.tiles += t;
        tiles += t;
        if (tiles > b->max_tiles)
            return bfail(b, MK_ERR_RANGE, "batch needs more than " + std::to_string(b->max_tiles) + " tiles");
    }
    H->n_items = n;

    hipStream_t st = (hipStream_t)stream;
    // A batch of one render request runs on the one-request launch path: VR in
    // two passes (the batched kernel marches every ray to its end in one), VR
    // and MIP sampling through the LDS address tables. The frame is the same;
    // the tables it reads were made current above, after the batches in
    // flight, so the path rebuilds nothing.
    if (n == 1 && items[0].kind == MK_ITEM_RENDER) {
        mk_context* c = b->ctxs[items[0].context];
        mk_status r = mk_render_async(c, &items[0].render, dev_arena + items[0].out_offset, stream);
        if (r) return bfail(b, r, std::string("item 0: ") + c->err);
        B_HIP(b, hipEventRecord(s.done, st));
        s.busy = true;
        return MK_OK;
    }

    B_HIP(b, hipGraphExecMemcpyNodeSetParams1D(s.exec, s.copy, s.d, s.h, b->lay.bytes(n), hipMemcpyHostToDevice));
    for (int v = 0; v < V_COUNT; ++v) {
        const bool used = H->var[v].tiles > 0;
        if (used) {
            hipKernelNodeParams kp = kernel_params(s, v, H->var[v].tiles);
            B_HIP(b, hipGraphExecKernelNodeSetParams(s.exec, s.kern[v], &kp));
        }
        if (used != s.enabled[v]) {
            B_HIP(b, hipGraphNodeSetEnabled(s.exec, s.kern[v], used ? 1u : 0u));
            s.enabled[v] = used;
        }
    }
    for (mk_context* c : b->ctxs) {
        mk_status r = mk_order_after_context(c, st);
        if (r) return bfail(b, r, c->err);
    }
    B_HIP(b, hipGraphLaunch(s.exec, st));
    B_HIP(b, hipEventRecord(s.done, st));
    s.busy = true;
    return MK_OK;
}
