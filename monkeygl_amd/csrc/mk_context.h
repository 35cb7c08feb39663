// mk_context.h -- host-side state of one render context (internal).
//
// One context replaces the reference's CuDataManager (core/CuDataManager.h:42-83):
// it owns the resident volume, mask and LUTs on one device and one HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/monkeygl_hip.h"
#include "mk_device.h"

// Per-brick statistics used for empty-space skipping. vmin/vmax cover the
// brick plus a one-voxel apron on every side (the trilinear footprint of any
// sample whose truncated index lies in the brick); labels covers the brick's
// own voxels (bit l = label l present, bit 64 in lab_hi).
struct MkBrickStats {
    int16_t* vmin = nullptr;
    int16_t* vmax = nullptr;
    uint64_t* lab_lo = nullptr;
    uint8_t* lab_hi = nullptr;
    uint8_t* blabel = nullptr;  // the brick's only label, 0xFF when it holds several
};

// LUT layout on the device: all labels' RGBA entries packed as float4, label l
// at [off[l], off[l]+len[l]); nzp holds, per label, len+1 prefix counts of
// entries with nonzero alpha (at offset off[l] + l).
struct MkLutTable {
    int off[MK_NUM_LABELS];
    int len[MK_NUM_LABELS];
};

// Loop trips of the first VR pass after which a ray moves to the second (the
// rays of its last waves move earlier), chosen per frame unless set: a frame
// whose waves all fit on the device at once (few rays, idle SIMDs) hands off
// early, one of up to one and a half such rounds after a middle budget, a
// larger one late.
constexpr int kVrTripBudgetAuto = -1;
constexpr int kVrTripBudgetSmall = 64;
constexpr int kVrTripBudgetMid = 256;
constexpr int kVrTripBudgetLarge = 1024;
constexpr int kVrResidentWavesPerCu = 12;   // 4 SIMDs x 3 waves of the first pass

// Skip tables of one render type. occ/occ2: brick / macro cell (4x4x4 bricks)
// may contribute (VR, Surface); mipb/mipb2: upper bound of the windowed value
// (MIP); bskip: per brick, what a ray entering it needs in one 16-bit load
// (MK_BSKIP_*: its only label, or 0xFF when mixed; occ; its macro cell's occ2).
// Valid for the (window/alpha palette, LUT version, volume version) they were
// built from.
#define MK_BSKIP_LABEL 0x00FFu
#define MK_BSKIP_OCC 0x0100u
#define MK_BSKIP_MACRO 0x0200u
struct MkSkipTables {
    uint8_t* occ = nullptr;
    float* mipb = nullptr;
    uint8_t* occ2 = nullptr;
    float* mipb2 = nullptr;
    uint16_t* bskip = nullptr;
    bool valid = false;
    uint32_t lut_version = 0;
    uint32_t vol_version = 0;
    mk_alpha_wwwl aw[MK_NUM_LABELS];
};

struct mk_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // volume
    int nx = 0, ny = 0, nz = 0;
    int nbx = 0, nby = 0, nbz = 0;
    int16_t* d_vox = nullptr;
    uint8_t* d_mask = nullptr;
    MkBrickStats stats;
    uint64_t used_lo = 1;   // labels present in the mask (label 0 without a mask)
    bool used_hi = false;

    // transfer functions
    std::vector<float> h_lut[MK_NUM_LABELS];
    bool lut_dirty = true;
    float4* d_lut = nullptr;
    int* d_lutnz = nullptr;
    size_t lut_cap = 0, lutnz_cap = 0;
    MkLutTable lut_tab{};
    uint32_t lut_version = 1;

    // per-frame skip tables, one set per render type, cached on their inputs
    MkSkipTables skip[3];

    // VR second pass: queue of the rays still marching after vr_budget loop
    // trips of the first pass (0: one pass); renders on other streams are
    // ordered after the last second pass (on cont_stream)
    int vr_budget = kVrTripBudgetAuto;
    int n_cu = 0;   // compute units of the device (queried once)
    void* d_cont = nullptr;
    size_t cont_bytes = 0;
    hipStream_t cont_stream = nullptr;
    bool cont_used = false;
    bool cont_clean = false;  // the queue counters are zero for the next frame
    uint32_t vol_version = 0;

    // exact float step lattice a_{k+1} = a_k + L1 from a_0 = 0
    float* d_steps = nullptr;
    size_t steps_cap = 0;
    float steps_l1 = -1.0f;
    int steps_n = 0;

    // marching cubes: per-block counts/offsets and the total (the triangle
    // table is per device, set up at context creation); calls on different
    // streams share this scratch, so a call waits on the device for the last
    // one (mc_done)
    void* d_mc_scratch = nullptr;
    size_t mc_scratch_cap = 0;
    void* d_mc_total = nullptr;
    hipEvent_t mc_done = nullptr;
    hipStream_t mc_stream = nullptr;
    bool mc_used = false;

    // Work enqueued on streams other than the context's that reads the
    // context's device data (volume, mask, statistics, LUTs, skip tables, step
    // lattice, orientation figure): one event per stream, recorded after each
    // such launch. Before the context rewrites or frees any of that data, its
    // own stream waits for every one of them (mk_wait_readers).
    struct Reader {
        hipStream_t stream;
        hipEvent_t done;
    };
    std::vector<Reader> readers;
    // orders a caller's stream after the context stream (mk_order_after_context)
    hipEvent_t order_ev = nullptr;

    // orientation figure of the VR overlay: normalised facets (9 floats each)
    float* d_orient = nullptr;
    int orient_n = 0;

    // annotation glyphs, index font * 256 + byte: width, height (-1 when the
    // byte has no glyph in that font) and mask bytes, packed into one device
    // atlas (glyph_off) when a draw finds them changed
    struct Glyph {
        int w = -1, h = -1;
        std::vector<uint8_t> mask;
    };
    std::vector<Glyph> glyphs;
    std::vector<size_t> glyph_off;
    bool glyphs_dirty = false;
    uint8_t* d_glyphs = nullptr;
    size_t glyphs_cap = 0;

    // staging for synchronous host-output calls
    void* d_stage = nullptr;
    size_t stage_cap = 0;
    void* d_stage2 = nullptr;
    size_t stage2_cap = 0;

    size_t nbricks() const { return (size_t)nbx * nby * nbz; }
    MkVolView view() const {
        MkVolView v;
        v.vox = d_vox;
        v.mask = d_mask;
        v.nx = nx; v.ny = ny; v.nz = nz;
        v.nbx = nbx; v.nby = nby; v.nbz = nbz;
        return v;
    }
};

// error helpers shared by the translation units
mk_status mk_fail(mk_context* ctx, mk_status st, const char* fmt, ...);
mk_status mk_hip_fail(mk_context* ctx, hipError_t e, const char* what);
mk_status mk_ensure_buffer(mk_context* ctx, void** buf, size_t* cap, size_t bytes);
mk_status mk_upload_luts(mk_context* ctx);
// Makes stream s wait for the work queued so far on the context's own stream
// (skip tables, lattice and table uploads); no-op when s is that stream.
mk_status mk_order_after_context(mk_context* ctx, hipStream_t s);
// Records that work just enqueued on stream s reads the context's device
// data; no-op when s is the context's stream (its own work is in order).
mk_status mk_note_reader(mk_context* ctx, hipStream_t s);
// Makes stream s wait for the work recorded so far by mk_note_reader on
// stream r (nothing when r's record was drained, which synchronized it).
mk_status mk_order_after_reader(mk_context* ctx, hipStream_t r, hipStream_t s);
// Makes the context's stream wait for all work recorded by mk_note_reader, so
// that what it enqueues next may rewrite that data; a synchronize of the
// context stream after this call also waits for those readers.
mk_status mk_wait_readers(mk_context* ctx);
// builds (on the context stream) the skip tables of render_type for the given
// palette unless the cached ones match; the LUTs must be uploaded
mk_status mk_ensure_skip_tables(mk_context* ctx, int render_type, const mk_alpha_wwwl* aw);
// uploads the float lattice a_{k+1} = a_k + l1 (a_0 = 0) up to past 1.732
mk_status mk_ensure_steps(mk_context* ctx, float l1);
// Argument checks of the one-request entry points, shared with the batched
// path so that a request gets the same status alone and in a batch.
mk_status mk_validate_render(mk_context* ctx, const mk_render_params* p);
mk_status mk_validate_plane(mk_context* ctx, const mk_plane_params* p);
mk_status mk_validate_cpr(mk_context* ctx, int width, int height, int sampling);
// once per device, with the device current: the marching-cubes triangle table
// and the device address of the allocation-free entry point's scratch
mk_status mk_mc_device_init(int device);

#define MK_HIP(ctx, call)                                          \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return mk_hip_fail((ctx), _e, #call); \
    } while (0)
