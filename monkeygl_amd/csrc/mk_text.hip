// mk_text.hip -- annotation text drawn into VR images.
//
// The reference's VR provider draws, after the image layer and before the
// CPR line (core/VRProvider.cpp:30-36), every annotation of the VR plane with
// AnnotationUtils::Textout2Image (core/AnnotationLayer.cpp:41-61,
// core/AnnotationUtils.cpp:419-464): the text's glyphs side by side from a
// start column (moved left by the text's width, or half of it, for right or
// centred text, AnnotationUtils::GetSize at 466-475), each glyph pixel with a
// nonzero mask byte m set to m * colour per channel. Glyph masks are the
// caller's data (mk_set_glyph); the reference compiles its own into its
// library (AnnotationUtils::Init, 42-376).
//
// The host lays the text out (every glyph's start column) exactly as the
// reference's loop does; one workgroup then draws the glyphs in order, each
// with all its threads, and waits between glyphs, since a byte without a
// glyph moves the cursor one pixel left and lets the next glyph overlap the
// previous one's last column, which the later glyph wins.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "mk_context.h"

namespace {

constexpr int kFonts = 3;
constexpr int kTextChunk = 96;     // glyphs per launch (kernel-argument block)
constexpr int kMaxGlyphSide = 1024;

struct TextChunk {
    int n;
    int y0;
    float color[3];
    int off[kTextChunk];     // glyph mask offset in the atlas
    int x0[kTextChunk];      // start column
    short w[kTextChunk];
    short h[kTextChunk];
};

// m * c in float, stored as the reference's float-to-unsigned-char
// conversion leaves it on the host: truncated toward zero, low 8 bits
__device__ inline uint8_t text_channel(uint8_t m, float c) {
    float v = (float)m * c;
    v = fminf(fmaxf(v, -2147483648.0f), 2147483520.0f);   // NaN -> the lower bound
    return (uint8_t)(int)v;
}

__global__ void __launch_bounds__(256) k_text(const uint8_t* __restrict__ atlas, TextChunk T, uint8_t* __restrict__ rgb,
                                              int width, int height) {
    for (int c = 0; c < T.n; ++c) {
        const int w = T.w[c], h = T.h[c], x0 = T.x0[c];
        const uint8_t* m = atlas + T.off[c];
        for (int i = threadIdx.x; i < w * h; i += blockDim.x) {
            const int gy = i / w, gx = i - gy * w;
            const int xx = x0 + gx, yy = T.y0 + gy;
            if (xx < 0 || xx >= width || yy < 0 || yy >= height) continue;
            const uint8_t v = m[i];
            if (v == 0) continue;
            uint8_t* px = rgb + 3 * ((size_t)yy * width + xx);
            px[0] = text_channel(v, T.color[0]);
            px[1] = text_channel(v, T.color[1]);
            px[2] = text_channel(v, T.color[2]);
        }
        __syncthreads();
    }
}

mk_status upload_glyphs(mk_context* ctx) {
    if (!ctx->glyphs_dirty && ctx->d_glyphs) return MK_OK;
    // draws on callers' streams may still read the current atlas
    mk_status st = mk_wait_readers(ctx);
    if (st) return st;
    size_t total = 0;
    ctx->glyph_off.assign(ctx->glyphs.size(), 0);
    for (size_t i = 0; i < ctx->glyphs.size(); ++i) {
        ctx->glyph_off[i] = total;
        total += ctx->glyphs[i].mask.size();
    }
    std::vector<uint8_t> packed(total ? total : 1, 0);
    for (size_t i = 0; i < ctx->glyphs.size(); ++i)
        if (!ctx->glyphs[i].mask.empty())
            memcpy(packed.data() + ctx->glyph_off[i], ctx->glyphs[i].mask.data(), ctx->glyphs[i].mask.size());
    st = mk_ensure_buffer(ctx, (void**)&ctx->d_glyphs, &ctx->glyphs_cap, packed.size());
    if (st) return st;
    MK_HIP(ctx, hipMemcpyAsync(ctx->d_glyphs, packed.data(), packed.size(), hipMemcpyHostToDevice, ctx->stream));
    // the packed vector dies at return
    MK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->glyphs_dirty = false;
    return MK_OK;
}

}  // namespace

extern "C" mk_status mk_set_glyph(mk_context* ctx, int font_size, int code, int width, int height,
                                  const uint8_t* mask) {
    if (!ctx) return MK_ERR_ARG;
    if (font_size < MK_FONT_SMALL || font_size > MK_FONT_BIG)
        return mk_fail(ctx, MK_ERR_ARG, "mk_set_glyph: font size %d", font_size);
    if (code < 0 || code > 255) return mk_fail(ctx, MK_ERR_ARG, "mk_set_glyph: code %d is not a byte", code);
    if (ctx->glyphs.empty()) ctx->glyphs.resize(kFonts * 256);
    mk_context::Glyph& g = ctx->glyphs[(size_t)font_size * 256 + code];
    if (width == 0) {
        g = mk_context::Glyph();
    } else {
        if (width < 1 || width > kMaxGlyphSide || height < 1 || height > kMaxGlyphSide || !mask)
            return mk_fail(ctx, MK_ERR_ARG, "mk_set_glyph: a %dx%d glyph (sides 1..%d, mask needed)", width, height,
                           kMaxGlyphSide);
        g.w = width;
        g.h = height;
        g.mask.assign(mask, mask + (size_t)width * height);
    }
    ctx->glyphs_dirty = true;
    return MK_OK;
}

extern "C" mk_status mk_draw_text_async(mk_context* ctx, const char* text, int x, int y, int font_size, int format,
                                        const float* color3, uint8_t* dev_rgb, int width, int height, void* stream) {
    if (!ctx) return MK_ERR_ARG;
    if (!text || !color3 || !dev_rgb || width <= 0 || height <= 0)
        return mk_fail(ctx, MK_ERR_ARG, "mk_draw_text_async: bad arguments");
    if (font_size < MK_FONT_NONE || font_size > MK_FONT_BIG)
        return mk_fail(ctx, MK_ERR_ARG, "mk_draw_text_async: font size %d", font_size);
    if (format < MK_ANNO_NONE || format > MK_ANNO_CENTER)
        return mk_fail(ctx, MK_ERR_ARG, "mk_draw_text_async: format %d", format);
    // Textout2Image's early returns
    if (width < 10 || height < 10 || format == MK_ANNO_NONE) return MK_OK;
    const size_t n = strlen(text);
    // glyph sizes per byte (the reference's default mask is -1 x -1)
    std::vector<int> gw(n, -1), gh(n, -1), gi(n, -1);
    long long total = 0;
    for (size_t i = 0; i < n; ++i) {
        if (font_size != MK_FONT_NONE && !ctx->glyphs.empty()) {
            const int k = font_size * 256 + (unsigned char)text[i];
            if (ctx->glyphs[k].w > 0) {
                gw[i] = ctx->glyphs[k].w;
                gh[i] = ctx->glyphs[k].h;
                gi[i] = k;
            }
        }
        total += gw[i];
    }
    long long xs = x;
    if (format == MK_ANNO_CENTER) xs -= total / 2;   // C++ int division, toward zero
    else if (format == MK_ANNO_RIGHT) xs -= total;
    MK_HIP(ctx, hipSetDevice(ctx->device));
    mk_status st = upload_glyphs(ctx);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    st = mk_order_after_context(ctx, s);
    if (st) return st;
    TextChunk T;
    memset(&T, 0, sizeof(T));
    T.y0 = y;
    for (int c = 0; c < 3; ++c) T.color[c] = color3[c];
    bool launched = false;
    for (size_t i = 0; i <= n; ++i) {
        const bool flush = i == n || T.n == kTextChunk;
        if (flush && T.n > 0) {
            hipLaunchKernelGGL(k_text, dim3(1), dim3(256), 0, s, ctx->d_glyphs, T, dev_rgb, width, height);
            MK_HIP(ctx, hipGetLastError());
            launched = true;
            T.n = 0;
        }
        if (i == n) break;
        // glyphs entirely left or right of the image draw nothing
        if (gi[i] >= 0 && xs + gw[i] > 0 && xs < width && (long long)y + gh[i] > 0 && y < height) {
            T.off[T.n] = (int)ctx->glyph_off[(size_t)gi[i]];
            T.x0[T.n] = (int)xs;
            T.w[T.n] = (short)gw[i];
            T.h[T.n] = (short)gh[i];
            ++T.n;
        }
        xs += gw[i];
    }
    return launched ? mk_note_reader(ctx, s) : MK_OK;
}
