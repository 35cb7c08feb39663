/*
 * monkeygl_hip.h -- C ABI of the MI355X (gfx950) render path.
 *
 * Plain C, POD structs, host or device pointers plus sizes. No HIP or torch
 * types appear in any signature, so the library can be bound from C++, from
 * ctypes or from any FFI.
 *
 * Every entry point below replaces one extern "C" launcher or device-data
 * call of the reference (MonkeyGL, core/kernel.cu and core/CuDataManager.cu);
 * the replaced symbol is cited beside each declaration. Texture objects and
 * cudaArrays of the reference are replaced by an opaque context that owns a
 * bricked int16 volume, a bricked uint8 label mask, per-label transfer
 * function LUTs and per-brick statistics, all resident in HBM.
 *
 * Error convention: every call returns mk_status (MK_OK == 0). A failing call
 * leaves a message retrievable with mk_last_error(ctx). Nothing ever exits
 * the process (the reference calls exit() from checkCudaErrors).
 *
 * Threading: a context is used by one host thread at a time. Synchronous
 * calls run on the context's own stream. Every call that takes a `stream`
 * argument enqueues its kernels on that hipStream_t (NULL = HIP's default
 * stream, as everywhere in HIP) after the context's pending preparation work
 * (skip tables, uploads), and returns without waiting. A later call that
 * rewrites or frees context data which earlier work on a caller's stream
 * still reads (LUTs, per-render-type skip tables, the step lattice, the
 * volume, the mask, the orientation figure) waits for that work on the device
 * first, so frames may be enqueued back to back with different palettes,
 * LUTs or step sizes. Marching-cubes calls of one context share its scratch
 * and run one after another on the device. Different contexts are
 * independent.
 */
#ifndef MONKEYGL_HIP_H
#define MONKEYGL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* core/Defines.h:34 MAXOBJECTCOUNT; labels 0..MK_MAX_OBJECTS are valid. */
#define MK_MAX_OBJECTS 64
#define MK_NUM_LABELS (MK_MAX_OBJECTS + 1)

/* Largest volume dimension, per axis (the reference sets none); a volume also
 * holds at most 2^31 voxels (4 GiB of int16) once padded to whole 8^3 bricks. */
#define MK_MAX_DIM 131072

typedef enum {
    MK_OK = 0,
    MK_ERR_ARG = 1,      /* null pointer, bad size, bad enum            */
    MK_ERR_STATE = 2,    /* e.g. render before any volume was uploaded  */
    MK_ERR_HIP = 3,      /* a HIP runtime call failed                   */
    MK_ERR_NOMEM = 4,    /* device allocation failed                    */
    MK_ERR_RANGE = 5     /* data out of range (label > MK_MAX_OBJECTS)  */
} mk_status;

/* core/Defines.h RenderType (RenderTypeVR, RenderTypeMIP, RenderTypeSurface) */
typedef enum { MK_RENDER_VR = 0, MK_RENDER_MIP = 1, MK_RENDER_SURFACE = 2 } mk_render_type;

/* core/Defines.h MPRType (MPRTypeAverage, MPRTypeMIP, MPRTypeMinIP) */
typedef enum { MK_MPR_AVERAGE = 0, MK_MPR_MIP = 1, MK_MPR_MINIP = 2 } mk_mpr_type;

/* Volume sampling for MPR/CPR. LINEAR is the reference's only mode
 * (core/CuDataManager.cu:64). NEAREST has no reference counterpart: it reads
 * the voxel floor(f * dim) (clamped) and is bit-exact across host and device. */
typedef enum { MK_SAMPLE_LINEAR = 0, MK_SAMPLE_NEAREST = 1 } mk_sampling;

/* core/Defines.h AlphaAndWWWL: per-label opacity cap and window. */
typedef struct { float alpha, ww, wl; } mk_alpha_wwwl;

/* core/Defines.h VOI: inclusive voxel index box. */
typedef struct { int left, right, posterior, anterior, head, foot; } mk_voi;

/*
 * Arguments of one VR / MIP / Surface frame: the by-value arguments of
 * cu_render (core/kernel.cu:471-492) minus the texture handles, which the
 * context owns. transform is the row-major float[9] that the reference copies
 * into float3x3 (core/CuDataManager.cpp:85-88).
 *
 * step_scale and exit_alpha parametrise the marcher: the reference rule is
 * step_scale = 1 (base step 1/depth, core/kernel.cu:169) and
 * exit_alpha = 0.995 (core/kernel.cu:248).
 */
typedef struct {
    int width;
    int height;
    int render_type;           /* mk_render_type */
    int invert_z;              /* VolumeInfo::Need2InvertZ */
    float max_len_spacing[3];  /* DataManager::InitCommon f3maxLenSpacing */
    float spacing[3];
    float spacing_voxel[3];    /* 1/dim per axis */
    float x_translate;
    float y_translate;
    float scale;
    float transform[9];
    mk_voi voi;
    float background[4];       /* RGBA background, core/kernel.cu:500 */
    mk_alpha_wwwl alpha_wwwl[MK_NUM_LABELS];
    float step_scale;
    float exit_alpha;
} mk_render_params;

/*
 * Arguments of one slab-MPR image: cu_renderPlane_{MIP,MinIP,Average}
 * (core/kernel.cu:616-631, 706-721, 795-810) minus the texture handle, plus
 * mpr_type (which launcher) and sampling (new: nearest).
 */
typedef struct {
    int width;
    int height;
    int mpr_type;        /* mk_mpr_type */
    int sampling;        /* mk_sampling */
    int invert_z;
    float spacing[3];
    float dir_h[3];
    float dir_v[3];
    float dir_n[3];
    float pt_left_top[3];
    float pixel_spacing;
    float half_num;
} mk_plane_params;

/* One marching-cubes triangle: three vertices, x,y,z each, voxel units. */
typedef struct { float v[9]; } mk_triangle;

typedef struct mk_context mk_context;

/* ---- device and context ------------------------------------------------ */

/* DeviceInfo::GetCount (core/DeviceInfo.cpp:91-97). */
int mk_device_count(void);

/* DeviceInfo + CuDataManager construction (core/DeviceInfo.cpp:28-70,
 * core/CuDataManager.cpp:34-44). device < 0 picks the device with the most
 * memory, as the reference does. */
mk_status mk_context_create(int device, mk_context** out);

/* CuDataManager::~CuDataManager (core/CuDataManager.cpp:46-68). */
void mk_context_destroy(mk_context* ctx);

/* Message of the last failing call on ctx (or of the last failing
 * mk_context_create when ctx is NULL). Never NULL. */
const char* mk_last_error(const mk_context* ctx);

/* The context's HIP stream as an opaque handle (hipStream_t). */
void* mk_context_stream(mk_context* ctx);

/* Blocks until all work queued on the context's stream is done. */
mk_status mk_synchronize(mk_context* ctx);

/* ---- device memory for bindings (no reference counterpart) ------------- */
/* For callers that hold no device-array library of their own (the pyMonkeyGL
 * module, plain C): memory on the context's device, and copies on the
 * context's stream that return when they are done (after the work queued
 * before them on that stream, e.g. an _async call given mk_context_stream).
 * mk_device_free waits for the context's stream before it frees. */
mk_status mk_device_alloc(mk_context* ctx, uint64_t bytes, void** dev_out);
mk_status mk_device_free(mk_context* ctx, void* dev);
mk_status mk_copy_to_host(mk_context* ctx, void* host, const void* dev, uint64_t bytes);
mk_status mk_copy_to_device(mk_context* ctx, void* dev, const void* host, uint64_t bytes);

/* ---- resident data ----------------------------------------------------- */

/* cu_copyVolumeData (core/CuDataManager.cu:28-73). host is int16, x fastest
 * (index z*nx*ny + y*nx + x), as the reference's make_cudaPitchedPtr layout.
 * Builds 8^3 bricks and per-brick min/max on the device. Drops any mask.
 * MK_ERR_ARG for a dimension above MK_MAX_DIM or more than 2^31 padded voxels. */
mk_status mk_copy_volume_data(mk_context* ctx, const int16_t* host, int nx, int ny, int nz);

/* Same, from a device buffer of the same layout (already in HBM). */
mk_status mk_copy_volume_data_device(mk_context* ctx, const int16_t* dev, int nx, int ny, int nz);

/* cu_copyMaskData (core/CuDataManager.cu:75-120). host is uint8 labels in
 * the volume's layout and dims; every value must be <= MK_MAX_OBJECTS. */
mk_status mk_copy_mask_data(mk_context* ctx, const uint8_t* host, int nx, int ny, int nz);
mk_status mk_copy_mask_data_device(mk_context* ctx, const uint8_t* dev, int nx, int ny, int nz);

/* Removes the mask (reference: mask texture 0 => every sample has label 0). */
mk_status mk_clear_mask(mk_context* ctx);

/* Incremental object-mask edit on the device. Replaces the reference's host
 * merge and full re-upload: VolumeInfo::AddNewObjectMask / UpdateObjectMask
 * (core/VolumeInfo.cpp:94-133) followed by CuDataManager::CopyMaskData
 * (core/DataManager.cpp:280-335). src is a binary mask in the volume's
 * layout and dims (nonzero = inside). MK_MASK_ADD: voxels inside take label
 * (without a mask yet the mask starts at 0, where the reference leaves the
 * rest of a new mask uninitialised). MK_MASK_UPDATE: also voxels outside that
 * hold label become 0; MK_ERR_STATE without a mask. The brick label sets are
 * rebuilt in the same pass; label must be 0..MK_MAX_OBJECTS. Like the
 * upload calls, run it when no render of this context is in flight. */
typedef enum { MK_MASK_ADD = 0, MK_MASK_UPDATE = 1 } mk_mask_mode;
mk_status mk_merge_object_mask(mk_context* ctx, const uint8_t* host, int nx, int ny, int nz, int label, int mode);
mk_status mk_merge_object_mask_device(mk_context* ctx, const uint8_t* dev, int nx, int ny, int nz, int label,
                                      int mode);

/* The label mask back in the volume's layout (nx*ny*nz bytes) into host
 * memory: DataManager::GetMaskData (core/DataManager.h:70). MK_ERR_STATE
 * without a mask. */
mk_status mk_get_mask_data(mk_context* ctx, uint8_t* host);

/* The image of HelloMonkey::GetOriginData_pngString (core/HelloMonkey.cpp:261-319):
 * axial slice z (clamped to the volume) as nx*ny int16, x fastest, with
 * -2048 where the mask holds 0 when there is a mask, into dev_out on stream. */
mk_status mk_origin_slice_async(mk_context* ctx, int z, int16_t* dev_out, void* stream);

/* cu_setTransferFunc (core/CuDataManager.cu:122-163) with the label of
 * CuDataManager::SetTransferFunction (core/CuDataManager.cpp:80-83).
 * rgba holds n RGBA float entries. A label with no LUT reads as all zero. */
mk_status mk_set_transfer_func(mk_context* ctx, int label, const float* rgba, int n);

/* ---- rendering ---------------------------------------------------------- */

/* cu_render (core/kernel.cu:471-568): renders into caller-owned host memory
 * of width*height*3 bytes (RGB24), synchronously. */
mk_status mk_render(mk_context* ctx, const mk_render_params* p, uint8_t* host_rgb);

/* Same frame into device memory (width*height*3 bytes) on the given stream;
 * returns after enqueueing. VR renders of one context share its queue of
 * second-pass rays (see mk_set_vr_trip_budget): one enqueued on another
 * stream than the previous one waits on the device for the previous one. */
mk_status mk_render_async(mk_context* ctx, const mk_render_params* p, uint8_t* dev_rgb, void* stream);

/* VR in two passes (no reference counterpart): the first pass marches every
 * ray until it ends, until `trips` loop trips, or until 95% of the first
 * pass's waves have finished (its last waves then hand off); the rays still
 * marching continue in a second pass that evaluates eight consecutive samples
 * of a ray at once and applies the reference's loop (core/kernel.cu:133-264)
 * to them in order. The frame is the same for every value; 0 renders in one
 * pass; -1 (the default) chooses per frame: 64 when all of the frame's waves
 * fit on the device at once, 256 when they fill it at most one and a half
 * times, else 1024. Frames whose queue would exceed 2 GiB render in one
 * pass. */
mk_status mk_set_vr_trip_budget(mk_context* ctx, int trips);

/* Measurement variant of mk_render_async (no reference counterpart): renders
 * the same frame without empty-space skipping and sets to 1 every voxel the
 * reference rule reads, in dense x-fastest byte maps of the volume's dims
 * (dev_touch_vol: volume voxels incl. gradient taps; dev_touch_mask: mask
 * voxels, may be NULL, ignored without a mask). The maps must be zeroed by the
 * caller. Their sums give the algorithmic bytes of the frame. */
mk_status mk_render_footprint(mk_context* ctx, const mk_render_params* p, uint8_t* dev_rgb,
                              uint8_t* dev_touch_vol, uint8_t* dev_touch_mask, void* stream);

/* cu_renderPlane_MIP / _MinIP / _Average (core/kernel.cu:616-837): int16
 * width*height output in caller-owned host memory, synchronously. */
mk_status mk_render_plane(mk_context* ctx, const mk_plane_params* p, int16_t* host_out);
mk_status mk_render_plane_async(mk_context* ctx, const mk_plane_params* p, int16_t* dev_out, void* stream);

/* cu_renderCPR (core/kernel.cu:875-918): row y samples points[y] + x*dirs[y]
 * (voxel units, 3 doubles each, height rows). Host pointers, synchronous. */
mk_status mk_render_cpr(mk_context* ctx, int width, int height, const double* points,
                        const double* dirs, int invert_z, int sampling, int16_t* host_out);
/* Same with device-resident points/dirs/output, on the given stream. */
mk_status mk_render_cpr_async(mk_context* ctx, int width, int height, const double* dev_points,
                              const double* dev_dirs, int invert_z, int sampling,
                              int16_t* dev_out, void* stream);

/* Marching cubes over the resident volume (no reference counterpart: the
 * reference carries only the triangle table, core/MarchingCube.cpp:35-292).
 * box = {x0, y0, z0, x1, y1, z1}, voxel indices, half-open; cubes have their
 * min corner in [x0, x1-1) x [y0, y1-1) x [z0, z1-1) and are visited x
 * fastest. A corner is inside when its raw int16 value < iso. Triangles come
 * in cube order, then table order; vertices are in voxel units.
 * Writes min(total, capacity) triangles to host_out (may be NULL with
 * capacity 0 to count only) and the total to *n_triangles. Synchronous. */
mk_status mk_marching_cubes(mk_context* ctx, const int* box, float iso, mk_triangle* host_out,
                            uint64_t capacity, uint64_t* n_triangles);

/* Same into device memory on the given stream: up to capacity triangles to
 * dev_out and the total (uint64) to dev_count (may be NULL). */
mk_status mk_marching_cubes_async(mk_context* ctx, const int* box, float iso, mk_triangle* dev_out,
                                  uint64_t capacity, uint64_t* dev_count, void* stream);

/* Allocation-free variant of mk_marching_cubes_async. No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same kernels and the same triangle list and count; the
 * per-block counts, offsets and total live in a fixed 12 MiB device array per
 * GPU, set up by mk_context_create, instead of the context's grown scratch.
 * The box may hold at most 2^28 cubes (2^20 blocks of 256); larger boxes
 * return MK_ERR_ARG. The call allocates nothing, records no event and does
 * not order itself after the context's own stream: upload the volume before
 * the call and do not replace it while the call is in flight. Every call on
 * one GPU shares that GPU's array (nothing in it needs clearing), so calls
 * that overlap on one GPU, from any context or stream, race: run them one
 * after another. */
mk_status mk_marching_cubes_noalloc(mk_context* ctx, const int* box, float iso, mk_triangle* dev_out,
                                    uint64_t capacity, uint64_t* dev_count, void* stream);

/* ---- differentiable transfer function (no reference counterpart) -------- */
/* Fitting rule: the VR compositing of mk_render on the same float step
 * lattice (base step 1/depth * step_scale), with a fixed step (no
 * refinement), every sample with alpha > 0 composited, no early exit, no
 * label caps (label 0's window from p->alpha_wwwl[0]; the mask is ignored),
 * sample alpha = min(1, opacity_scale * LUT_alpha) and the float RGB sum as
 * output. dev_lut holds n RGBA float entries (2 <= n <= 4096) in device
 * memory: the LUT is a parameter, not the context's label LUT. */

/* Forward: width*height*3 floats (RGB) into dev_rgb and, unless dev_rgb64 is
 * NULL, the same colours as width*height*3 doubles into dev_rgb64 (what the
 * adjoint takes to skip a march). */
mk_status mk_tf_forward(mk_context* ctx, const mk_render_params* p, const float* dev_lut, int n,
                        float opacity_scale, float* dev_rgb, double* dev_rgb64, void* stream);

/* Adjoint: given dL/dRGB (width*height*3 floats), ADDS dL/dLUT (n*4 doubles)
 * and dL/d(opacity_scale) (1 double) into the caller's zeroed or running
 * accumulators. Each ray is re-marched once for the gradients, un-compositing
 * the rest of the ray with a (1 - alpha) division from its final colour:
 * dev_rgb64 as mk_tf_forward wrote it for the same view, LUT and scale, or
 * NULL, when a first march recomputes it. Nothing per sample is stored.
 * Accepts the same 2..4096 entries as the forward: up to 1365 entries each
 * workgroup sums its gradients in LDS first, above that the lanes add them to
 * the accumulators in device memory directly (same sums, in another order). */
mk_status mk_tf_backward(mk_context* ctx, const mk_render_params* p, const float* dev_lut, int n,
                         float opacity_scale, const float* dev_grad_rgb, const double* dev_rgb64,
                         double* dev_grad_lut, double* dev_grad_scale, void* stream);

/* ---- VR overlay: the orientation figure --------------------------------- */

/* The figure drawn into the corner of VR images (MarchingCube::LoadBinarySTL,
 * core/MarchingCube.cpp:306-365): n_facets facets of a binary STL body (the
 * file after its 84-byte header: 50 bytes per facet, normal and three
 * vertices as float32, a 2-byte attribute), normalised as the reference does:
 * centred on the bounding box, divided by the larger of the x and z extents,
 * z negated. n_facets = 0 removes the figure. */
mk_status mk_set_orientation_mesh(mk_context* ctx, const uint8_t* stl_facets, int n_facets);

/* Device workspace mk_merge_orientation_box_async needs for ctx's figure. */
uint64_t mk_orientation_workspace_bytes(mk_context* ctx);

/* VRImageLayer::MergeOrientationBox (core/VRImageLayer.cpp:104-194) on an
 * RGB24 device image: the figure seen through rotate9 (RenderInfo's rotate
 * matrix, 9 floats, row-major), rasterised at 512x512 as
 * Methods::FillHoleInImage_Ch1 does (core/Methods.cpp:308-368), every 8th
 * pixel lighting the bottom-right 64x64 corner. No-op when width or height is
 * at most 64 or no figure is set. The workspace (8-byte aligned) holds the
 * per-call state, so calls with their own workspaces may overlap. */
mk_status mk_merge_orientation_box_async(mk_context* ctx, const float* rotate9, uint8_t* dev_rgb, int width,
                                         int height, void* dev_workspace, uint64_t workspace_bytes, void* stream);

/* CPRLineLayer::GetRGBData (core/CPRLineLayer.cpp:40-65) on an RGB24 device
 * image: the polyline through n points (dev_xy: 2n floats, VR image
 * coordinates, i.e. the CPR line's voxel points through
 * DataManager::TransferVoxel2ImageInVR, core/DataManager.cpp:236-258, done by
 * the caller), each segment drawn as Methods::DrawLineInImage24Bit of width 1
 * (core/Methods.cpp:183-200) in color3 (RGB in 0..1, stored as int(c * 255)).
 * Fewer than 2 points draw nothing. */
mk_status mk_draw_polyline_async(mk_context* ctx, const float* dev_xy, int n, const float* color3, uint8_t* dev_rgb,
                                 int width, int height, void* stream);

/* ---- VR overlay: annotation text ---------------------------------------- */

/* core/Defines.h FontSize and AnnotationFormat. */
typedef enum { MK_FONT_NONE = -1, MK_FONT_SMALL = 0, MK_FONT_MIDDLE = 1, MK_FONT_BIG = 2 } mk_font_size;
typedef enum { MK_ANNO_NONE = -1, MK_ANNO_LEFT = 0, MK_ANNO_RIGHT = 1, MK_ANNO_CENTER = 2 } mk_anno_format;

/* The coverage mask of one character (a byte) in one font size: width x
 * height bytes, row-major, as AnnotationUtils keeps its character masks
 * (core/AnnotationUtils.cpp:42-376, Mask in core/AnnotationUtils.h:31-46).
 * The reference compiles its masks into its library; here the caller
 * registers them (the same data, or any font). width and height are
 * 1..1024; width 0 removes the glyph. A byte without a glyph behaves as the
 * reference's default mask: width and height -1, so it draws nothing and
 * moves the text cursor one pixel left. */
mk_status mk_set_glyph(mk_context* ctx, int font_size, int code, int width, int height, const uint8_t* mask);

/* AnnotationUtils::Textout2Image (core/AnnotationUtils.cpp:419-464) with the
 * font size set first (AnnotationUtils::SetFontSize), as
 * AnnotationLayer::GetRGBData draws each annotation (core/AnnotationLayer.cpp:
 * 41-61), on an RGB24 device image on the given stream. text is a
 * NUL-terminated byte string, one glyph per byte; (x, y) is the top-left of
 * the text, moved left by its width (the sum of its glyph widths, -1 per
 * byte without one) for MK_ANNO_RIGHT and by half of it (integer division)
 * for MK_ANNO_CENTER. Every glyph pixel inside the image whose mask byte m is
 * nonzero becomes m * color3 per channel, computed in float and stored
 * truncated toward zero, modulo 256 (colour components in [0, 1] keep it in
 * range). Later glyphs overwrite earlier ones where they overlap. Nothing is
 * drawn when width or height is below 10, for MK_ANNO_NONE, or for
 * MK_FONT_NONE. */
mk_status mk_draw_text_async(mk_context* ctx, const char* text, int x, int y, int font_size, int format,
                             const float* color3, uint8_t* dev_rgb, int width, int height, void* stream);

/* ---- output encoding: PNG + base64 on the device ------------------------- */
/* Replaces the host encoding of served images: fpng_encode_image_to_memory
 * of the RGB VR frame (core/HelloMonkey.cpp:333-365) and of an int16 plane
 * read as an RGBA image of width W/2 (core/HelloMonkey.cpp:212-259), then
 * Base64::Encode (core/Base64.hpp). The PNG differs in its bytes from fpng's
 * but decodes to the same pixels (8-bit RGB or RGBA, one IDAT chunk, zlib
 * stream with one deflate block coded with the image's own Huffman code).
 * Rows hold at most 24576 bytes (width * channels), height at most 65536. */

/* Upper bound of the encoded PNG size (0 for unsupported dims). */
uint64_t mk_png_max_bytes(int width, int height, int channels);
/* Device workspace the encoder needs (256-byte aligned). */
uint64_t mk_png_workspace_bytes(int width, int height, int channels);
/* Encodes width x height x channels (3 or 4) bytes at dev_pixels into dev_png
 * (capacity >= mk_png_max_bytes, 4-byte aligned) on the stream; the length is
 * written to *dev_png_len (device memory) when the stream reaches it. Errors
 * of the arguments are reported through mk_last_error(NULL). */
mk_status mk_png_encode_async(const uint8_t* dev_pixels, int width, int height, int channels, void* dev_workspace,
                              uint64_t workspace_bytes, uint8_t* dev_png, uint64_t png_capacity,
                              uint64_t* dev_png_len, void* stream);
/* 4 * ceil(n / 3). */
uint64_t mk_base64_max_bytes(uint64_t n);
/* Base64 (standard alphabet, '=' padding, no line breaks) of the
 * min(*dev_in_len, in_capacity) bytes at dev_in, with the length read on the
 * device, so it can follow mk_png_encode_async on the same stream. */
mk_status mk_base64_encode_async(const uint8_t* dev_in, const uint64_t* dev_in_len, uint64_t in_capacity,
                                 uint8_t* dev_out, uint64_t out_capacity, uint64_t* dev_out_len, void* stream);

/* ---- batched serving launches (no reference counterpart) ----------------- */
/* The render server batches requests over several resident volumes into one
 * launch. Outputs go to a caller-owned device arena at each item's offset:
 * RGB24 for RENDER items, int16 for PLANE and CPR items (even offsets). CPR
 * items read their height*3 points and height*3 directions (doubles, 8-byte
 * aligned offsets) from the same arena. */
typedef enum { MK_ITEM_RENDER = 0, MK_ITEM_PLANE = 1, MK_ITEM_CPR = 2 } mk_item_kind;

typedef struct {
    int kind;                    /* mk_item_kind */
    int context;                 /* index into the batcher's contexts */
    uint64_t out_offset;         /* byte offset of the output in the arena */
    mk_render_params render;     /* RENDER: as mk_render_async */
    mk_plane_params plane;       /* PLANE: as mk_render_plane_async */
    int cpr_width, cpr_height, cpr_invert_z, cpr_sampling;  /* CPR: as mk_render_cpr_async */
    uint64_t cpr_points_offset;
    uint64_t cpr_dirs_offset;
} mk_batch_item;

typedef struct mk_batcher mk_batcher;

/* A batcher over nctx contexts of one device with nslots launch slots. Each
 * slot holds a hipGraph built once: descriptor upload, then one kernel per
 * kernel variant (VR with/without mask and LDS LUT, MIP with/without mask,
 * Surface, MPR/CPR), each over its variant's 16x16 tiles. Batches may hold up
 * to max_items items and max_tiles tiles in all. */
mk_status mk_batcher_create(mk_context* const* ctxs, int nctx, int max_items, int max_tiles, int nslots,
                            mk_batcher** out);
void mk_batcher_destroy(mk_batcher* b);
/* Last error of b, or of the last failing mk_batcher_create when b is NULL. */
const char* mk_batcher_last_error(const mk_batcher* b);

/* Launches a batch in a free slot on the given stream (NULL: the HIP default
 * stream) and returns. A render item whose LUTs, skip tables or step lattice
 * must be rebuilt first waits for every batch in flight. Within one batch,
 * the render items of one context share one step_scale, and those of one
 * context and render type one alpha_wwwl table (MK_ERR_ARG otherwise). A
 * batch of one render item is rendered as mk_render_async renders it (VR in
 * two passes), same image, after the same checks. A context used by a batcher
 * must not be rendered through the one-request entry points while its
 * batches run. */
mk_status mk_batch_submit(mk_batcher* b, int slot, const mk_batch_item* items, int n, uint8_t* dev_arena,
                          uint64_t arena_bytes, void* stream);

/* 1 when the slot is free (its last batch finished), 0 while it runs, -1 on error. */
int mk_batch_query(mk_batcher* b, int slot);

/* Blocks until the slot's last batch has finished. */
mk_status mk_batch_wait(mk_batcher* b, int slot);

#ifdef __cplusplus
}
#endif

#endif /* MONKEYGL_HIP_H */
