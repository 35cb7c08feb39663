// mk_encode.hip -- PNG and base64 encoding of rendered images on the GPU.
//
// The reference encodes every served image on the host: VR frames as RGB PNG
// and int16 MPR/CPR images as RGBA PNG of width W/2 (the int16 buffer read as
// bytes), both through fpng, then base64 (core/HelloMonkey.cpp:212-259,
// 333-413; core/Base64.hpp). Here the image stays in HBM and the PNG is built
// there; any PNG decoder returns the same pixels, which is the contract the
// client (a PNG decoder after atob) relies on.
//
// Format. One IDAT chunk holding a zlib stream with a single deflate block
// coded with Huffman codes built for the image (dynamic block). Each row gets
// the PNG filter (None, Sub, Up, Average, Paeth) with the smallest sum of
// |signed byte| (the usual heuristic); the filtered row is coded as runs of
// equal bytes: a run's first byte is a literal and the rest are distance-1
// copies in matches of up to 258 (a remainder below 3 stays literal). The
// literal/length code is a Huffman code of the image's symbol counts, limited
// to 12 bits (so no byte costs more than 12 bits); the one distance code has
// length 1; the code lengths are sent with a fixed 4-bit code-length code.
// Rows never reference each other, so every row is coded independently: its
// bit length follows from its symbol counts, an exclusive scan gives the
// offsets, then every row writes its bits at its offset. Adler-32 is combined
// from per-row sums, CRC-32 from per-block raw CRCs (CRC is linear: a block's
// contribution is moved past the following bytes by multiplying with x^(8n)
// mod P).
//
// Launch sequence (one stream, no host round trip): filter + symbol counts
// (one workgroup per row) -> code + scan (one workgroup: the Huffman code,
// row offsets, every byte around the row data) -> emit rows -> CRC
// (workgroups of 4 KiB; the last one to finish writes the CRC, IEND and the
// length). The encoded length is left in device memory, so base64 can follow
// on the same stream. Rows are limited to 24576 bytes: the emit kernel holds
// a filtered row and its bits in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "mk_context.h"

namespace {

constexpr int kMaxRowBytes = 24576;          // width * channels (row + its bits fit in LDS)
constexpr uint32_t kCrcPoly = 0xEDB88320u;   // reflected CRC-32 polynomial
constexpr uint32_t kAdlerMod = 65521u;
constexpr int kCrcBlock = 4096;              // bytes per CRC workgroup
constexpr uint64_t kHeadBytes = 8 + 25 + 8;  // signature, IHDR, IDAT length + type

constexpr int kNumLitLen = 286;     // literal/length symbols 0..285 (256 = end of block)
constexpr int kMaxCodeBits = 12;    // literal/length code length limit
constexpr int kHistStride = 288;    // uint16 counts per row histogram (16-byte rows)
// dynamic block header: BFINAL + BTYPE, HLIT, HDIST, HCLEN, 19 code-length
// code lengths, then 286 literal/length and 1 distance code lengths of 4 bits
constexpr int kHeaderBits = 3 + 5 + 5 + 4 + 19 * 3 + (kNumLitLen + 1) * 4;
constexpr uint32_t kBlockStartBit = (uint32_t)(kHeadBytes + 2) * 8;   // after the zlib header (bit 344)

__host__ __device__ inline uint32_t bit_reverse(uint32_t c, int n) {
    uint32_t r = 0;
    for (int i = 0; i < n; ++i) r |= ((c >> i) & 1u) << (n - 1 - i);
    return r;
}

// Length symbol (257..285), extra-bit count and base of a match of length L
// in 3..258 (RFC 1951 3.2.5).
__device__ inline void len_symbol(int L, int& sym, int& nextra, int& base) {
    if (L == 258) {
        sym = 285; nextra = 0; base = 258;
    } else if (L <= 10) {
        sym = 254 + L; nextra = 0; base = L;
    } else {
        // groups of four symbols with 1..5 extra bits, starting at 11, 19, 35, 67, 131
        int e = 1, start = 11, s0 = 265;
        while (L >= start + (4 << e)) { start += 4 << e; s0 += 4; ++e; }
        const int k = (L - start) >> e;
        sym = s0 + k; nextra = e; base = start + (k << e);
    }
}

// ---- GF(2) arithmetic for CRC combination (reflected representation) -------

__host__ __device__ inline uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
    // a * b mod P, bit 31 = x^0
    uint32_t p = 0;
    for (int i = 0; i < 32; ++i) {
        if (a & (0x80000000u >> i)) p ^= b;
        b = (b & 1u) ? (b >> 1) ^ kCrcPoly : b >> 1;
    }
    return p;
}

// x^(8 * 2^k) mod P for k < 48, built once on the host: x^(8 n) is then a
// product over the set bits of n.
struct PowTable {
    uint32_t p[48];
};

PowTable make_pow_table() {
    PowTable t;
    uint32_t sq = 0x00800000u;   // x^8
    for (int k = 0; k < 48; ++k) {
        t.p[k] = sq;
        sq = gf2_mulmod(sq, sq);
    }
    return t;
}

__device__ inline uint32_t x_pow_8n_tab(const PowTable& t, uint64_t n) {
    uint32_t r = 0x80000000u;
    for (int k = 0; n; ++k, n >>= 1)
        if (n & 1) r = gf2_mulmod(r, t.p[k]);
    return r;
}

__device__ inline uint32_t crc_update_raw(uint32_t c, uint8_t byte, const uint32_t* table) {
    return table[(c ^ byte) & 0xFFu] ^ (c >> 8);
}

// ---- kernels -----------------------------------------------------------------

struct PngWs {
    uint8_t* filt;      // H * (1 + stride)
    uint32_t* row_bits; // H
    uint64_t* row_s;    // H: sum of filtered bytes, mod 65521
    uint64_t* row_t;    // H: sum of (len - k) * byte, k = 0..len-1, mod 65521
    uint64_t* row_off;  // H: bit offset of each row
    uint64_t* totals;   // [0] end of the row bits, [1] zlib length, [2] png length, [3] adler, [4] CRC ticket
    uint32_t* crc_part; // per CRC block
    uint16_t* hist;     // H * kHistStride: each row's literal/length symbol counts
    uint32_t* extra;    // H: each row's extra and distance bits
    uint32_t* code;     // kNumLitLen: bit-reversed code | length << 24
};

// PNG filter per row; one workgroup per row.
__device__ inline int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

__device__ inline uint8_t filt_byte(int f, const uint8_t* cur, const uint8_t* prev, int i, int bpp) {
    const int x = cur[i];
    const int a = i >= bpp ? cur[i - bpp] : 0;
    const int b = prev ? prev[i] : 0;
    const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
    switch (f) {
    case 0: return (uint8_t)x;
    case 1: return (uint8_t)(x - a);
    case 2: return (uint8_t)(x - b);
    case 3: return (uint8_t)(x - ((a + b) >> 1));
    default: return (uint8_t)(x - paeth(a, b, c));
    }
}

// Workgroup-wide helpers for the row kernels (kRowThreads threads, 4 waves).
constexpr int kRowThreads = 256;
constexpr int kRowWaves = kRowThreads / 64;

// Exclusive prefix sum over the workgroup; total gets the sum of all.
__device__ inline uint32_t block_excl_sum(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int w = 0; w < kRowWaves; ++w) {
        const uint32_t x = s_w[w];
        if (w < wave) before += x;
        total += x;
    }
    __syncthreads();   // s_w may be reused
    return before + incl - v;
}

// Minimum of v over the threads after this one (INT_MAX for the last).
__device__ inline int block_min_after(int v, int* s_w) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;   // suffix minimum within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_down(incl, o, 64);
        if (lane + o < 64) incl = min(incl, y);
    }
    int after = __shfl_down(incl, 1, 64);
    if (lane == 63) after = INT_MAX;
    if (lane == 0) s_w[wave] = incl;
    __syncthreads();
    for (int w = wave + 1; w < kRowWaves; ++w) after = min(after, s_w[w]);
    __syncthreads();
    return after;
}

// Runs of one filtered row (n bytes in LDS), parsed by the whole workgroup.
// A run starts at byte 0 and wherever a byte differs from the previous one,
// and ends at the next start (or the row end). Thread t owns the bytes
// [t k, (t + 1) k), k = ceil(n / kRowThreads), and the runs starting there;
// the end of its last run is the first start owned by a later thread.
struct RowRuns {
    int first;   // the thread's first run start (INT_MAX: none)
    int next;    // the first start after the thread's bytes (n: none)
    int hi;      // end of the thread's bytes
};

// Every thread of the workgroup calls it (it holds barriers).
__device__ inline RowRuns row_runs(const uint8_t* row, int n, int* s_i) {
    const int k = (n + kRowThreads - 1) / kRowThreads;
    const int lo = min(n, (int)threadIdx.x * k), hi = min(n, lo + k);
    int first = INT_MAX;
    for (int i = lo; i < hi; ++i)
        if (i == 0 || row[i] != row[i - 1]) {
            first = i;
            break;
        }
    int next = block_min_after(first, s_i);
    if (next == INT_MAX) next = n;
    return RowRuns{first, next, hi};
}

// fn(value, length) for each run of the calling thread, in row order.
template <typename Fn>
__device__ inline void visit_runs(const uint8_t* row, const RowRuns& R, Fn fn) {
    if (R.first == INT_MAX) return;
    int st = R.first, v = row[st];
    for (int i = st + 1; i < R.hi; ++i) {
        const int b = row[i];
        if (b != v) {
            fn(v, i - st);
            st = i;
            v = b;
        }
    }
    fn(v, R.next - st);
}

// The symbols of a run of n equal bytes v: one literal, then n - 1 copies in
// matches of up to 258 (a remainder below 3 stays literal). lit(v, count),
// match(length) for each match.
template <typename Lit, typename Match>
__device__ inline void run_symbols(int v, int n, Lit lit, Match match) {
    lit(v, 1);
    int m = n - 1;
    while (m >= 258) {
        match(258);
        m -= 258;
    }
    if (m >= 3) match(m);
    else if (m > 0) lit(v, m);
}

// Filter, symbol counts and Adler partials of one row; one workgroup per
// row, the filtered row staged in LDS for the run parse.
__global__ void __launch_bounds__(kRowThreads) k_png_filter_hist(const uint8_t* __restrict__ img, int width, int height,
                                                          int ch, PngWs ws) {
    extern __shared__ uint8_t s_row[];
    const int r = blockIdx.x;
    const int stride = width * ch;
    const int n = stride + 1;
    const uint8_t* cur = img + (size_t)r * stride;
    const uint8_t* prev = r > 0 ? cur - stride : nullptr;
    __shared__ uint32_t s_sum[5][4];
    __shared__ uint64_t s_ad[2][4];
    __shared__ int s_best;
    uint32_t acc[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < stride; i += blockDim.x)
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const uint8_t v = filt_byte(f, cur, prev, i, ch);
            acc[f] += v < 128 ? v : 256u - v;
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        uint32_t v = acc[f];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) s_sum[f][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t best = ~0ull;
        int bf = 0;
        for (int f = 0; f < 5; ++f) {
            const uint64_t t = (uint64_t)s_sum[f][0] + s_sum[f][1] + s_sum[f][2] + s_sum[f][3];
            if (t < best) { best = t; bf = f; }
        }
        s_best = bf;
    }
    __syncthreads();
    const int f = s_best;
    uint8_t* out = ws.filt + (size_t)r * n;
    uint64_t sa = 0, sb = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const uint8_t b = i == 0 ? (uint8_t)f : filt_byte(f, cur, prev, i - 1, ch);
        s_row[i] = b;
        out[i] = b;
        sa += b;
        sb += (uint64_t)(n - i) * b;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sa += __shfl_down(sa, o, 64);
        sb += __shfl_down(sb, o, 64);
    }
    if (lane == 0) {
        s_ad[0][wave] = sa;
        s_ad[1][wave] = sb;
    }
    __shared__ uint32_t s_u[kRowWaves];
    __shared__ int s_i[kRowWaves];
    __shared__ uint32_t s_hist[kHistStride];
    for (int i = threadIdx.x; i < kHistStride; i += blockDim.x) s_hist[i] = 0u;
    __syncthreads();
    const RowRuns R = row_runs(s_row, n, s_i);
    uint32_t extra = 0;   // length extra bits and one distance bit per match
    visit_runs(s_row, R, [&](int v, int len) {
        run_symbols(
            v, len, [&](int b, int count) { atomicAdd(&s_hist[b], (uint32_t)count); },
            [&](int L) {
                int sym, ne, base;
                len_symbol(L, sym, ne, base);
                atomicAdd(&s_hist[sym], 1u);
                extra += (uint32_t)ne + 1u;
            });
    });
    uint32_t total_extra;
    (void)block_excl_sum(extra, s_u, total_extra);   // ends with a barrier: the counts are complete
    uint16_t* h = ws.hist + (size_t)r * kHistStride;
    for (int i = threadIdx.x; i < kHistStride; i += blockDim.x) h[i] = (uint16_t)s_hist[i];
    if (threadIdx.x == 0) {
        ws.extra[r] = total_extra;
        // the row's byte sum and weighted sum, reduced mod 65521 for the Adler-32 combination
        ws.row_s[r] = (s_ad[0][0] + s_ad[0][1] + s_ad[0][2] + s_ad[0][3]) % kAdlerMod;
        ws.row_t[r] = (s_ad[1][0] + s_ad[1][1] + s_ad[1][2] + s_ad[1][3]) % kAdlerMod;
    }
}

// One workgroup of kScanThreads: the image's literal/length Huffman code, the
// row bit lengths and offsets, Adler-32, and every byte of the PNG around the
// row data: signature, IHDR, IDAT length and type, zlib header and the
// dynamic block header before it; end-of-block code, padding and Adler-32
// after it. The output is not cleared beforehand: the emit kernel stores a
// row's interior words whole and ORs only the first and last word of each
// row, which this kernel zeroes, with the words of the end-of-block code.
//
// The code: symbol counts summed over the rows' histograms (plus one end of
// block), sorted by a bitonic sort, optimal lengths limited to kMaxCodeBits
// by package-merge (one parallel merge per level), canonical codes.
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads) k_huff_scan(int height, int stride, PngWs ws, uint8_t* png,
                                                            uint32_t ihdr_crc, int width, int ch) {
    __shared__ uint32_t s_freq[kNumLitLen];
    __shared__ unsigned long long s_key[512];
    __shared__ int s_npk[kMaxCodeBits + 1];        // packages per level
    __shared__ int s_cnt[kMaxCodeBits + 1];        // leaves chosen per level
    __shared__ uint16_t s_pos[kMaxCodeBits * kNumLitLen];   // leaf positions per level
    __shared__ uint32_t s_len2[kHistStride / 2];   // code lengths of symbol pairs (low | high << 16)
    __shared__ int s_wcnt[(kNumLitLen + 63) / 64][kMaxCodeBits + 1];
    __shared__ int s_blc[kMaxCodeBits + 2];
    __shared__ int s_next[kMaxCodeBits + 2];
    __shared__ uint8_t s_len[kNumLitLen];
    __shared__ uint32_t s_code[kNumLitLen];
    __shared__ uint32_t s_hdr[40];                 // bits 320.. of the PNG: bytes 40-42 and the block header
    __shared__ int s_m;
    // the package lists of levels 1..kMaxCodeBits-1 and the merged list, then
    // (once the code is built) the scan's arrays
    constexpr int kBig = (kMaxCodeBits - 1) * kNumLitLen + 2 * kNumLitLen > 3 * kScanThreads
                             ? (kMaxCodeBits - 1) * kNumLitLen + 2 * kNumLitLen
                             : 3 * kScanThreads;
    __shared__ unsigned long long s_big[kBig];
    unsigned long long* s_pk = s_big;                                   // [level - 1][kNumLitLen]
    unsigned long long* s_list = s_big + (kMaxCodeBits - 1) * kNumLitLen;   // 2 * kNumLitLen
    unsigned long long* s_part = s_big;
    unsigned long long* s_a = s_big + kScanThreads;
    unsigned long long* s_b = s_big + 2 * kScanThreads;
    const int t = threadIdx.x, nt = blockDim.x;

    // 1. symbol counts of the image
    for (int i = t; i < kNumLitLen; i += nt) s_freq[i] = 0u;
    for (int i = t; i < 40; i += nt) s_hdr[i] = 0u;
    __syncthreads();
    {
        // thread t sums symbols 4q..4q+3 (q = t % 72) over the rows of group
        // t / 72 (coalesced loads; a few atomics per symbol instead of one per row)
        constexpr int kQuads = kHistStride / 4;
        const int groups = nt / kQuads;
        if (t < groups * kQuads) {
            const int q = t % kQuads, g = t / kQuads;
            uint32_t c[4] = {0, 0, 0, 0};
            int r = g;
            for (; r + 3 * groups < height; r += 4 * groups) {
                uint2 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = ((const uint2*)(ws.hist + (size_t)(r + u * groups) * kHistStride))[q];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    c[0] += v[u].x & 0xFFFFu; c[1] += v[u].x >> 16;
                    c[2] += v[u].y & 0xFFFFu; c[3] += v[u].y >> 16;
                }
            }
            for (; r < height; r += groups) {
                const uint2 v = ((const uint2*)(ws.hist + (size_t)r * kHistStride))[q];
                c[0] += v.x & 0xFFFFu; c[1] += v.x >> 16;
                c[2] += v.y & 0xFFFFu; c[3] += v.y >> 16;
            }
            for (int u = 0; u < 4; ++u)
                if (4 * q + u < kNumLitLen && c[u]) atomicAdd(&s_freq[4 * q + u], c[u]);
        }
    }
    __syncthreads();
    if (t == 0) s_freq[256] = 1u;   // end of block
    __syncthreads();

    // 2. used symbols sorted by (count, symbol)
    for (int i = t; i < 512; i += nt)
        s_key[i] = (i < kNumLitLen && s_freq[i]) ? ((unsigned long long)s_freq[i] << 16 | (unsigned)i) : ~0ull;
    __syncthreads();
    for (int size = 2; size <= 512; size <<= 1)
        for (int st = size >> 1; st > 0; st >>= 1) {
            if (t < 512) {
                const int j = t ^ st;
                if (j > t) {
                    const unsigned long long a = s_key[t], b = s_key[j];
                    const bool up = (t & size) == 0;
                    if ((a > b) == up) {
                        s_key[t] = b;
                        s_key[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    if (t < 512 && s_key[t] != ~0ull && (t == 511 || s_key[t + 1] == ~0ull)) s_m = t + 1;
    __syncthreads();
    const int m = s_m;   // >= 2: the end of block and every row's filter byte

    // 3. optimal code lengths limited to kMaxCodeBits by package-merge: the
    // list of level kMaxCodeBits is the leaves; the list of level l < that
    // merges the leaves with the pairs ("packages") of the list of level
    // l + 1, leaves first on equal weights. The first 2m - 2 items of level 1
    // are chosen; a chosen package at level l chooses the first two items of
    // level l + 1 per package. A leaf's code length is the number of levels
    // at which it is chosen, and the leaves chosen at a level are a prefix of
    // the sorted leaves, so only their count per level is needed.
    auto leaf_w = [&](int i) { return s_key[i] >> 16; };
    int src_len = m;   // the list of level l + 1 (the leaves at the deepest level)
    for (int l = kMaxCodeBits - 1; l >= 1; --l) {
        unsigned long long* pk = s_pk + (size_t)(l - 1) * kNumLitLen;
        const int np = src_len / 2;
        for (int j = t; j < np; j += nt)
            pk[j] = l == kMaxCodeBits - 1 ? leaf_w(2 * j) + leaf_w(2 * j + 1) : s_list[2 * j] + s_list[2 * j + 1];
        if (t == 0) s_npk[l] = np;
        __syncthreads();
        // merged list of level l
        for (int i = t; i < m; i += nt) {
            const unsigned long long w = leaf_w(i);
            int lo = 0, hi = np;   // packages lighter than the leaf
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (pk[mid] < w) lo = mid + 1;
                else hi = mid;
            }
            s_list[i + lo] = w;
        }
        for (int j = t; j < np; j += nt) {
            const unsigned long long w = pk[j];
            int lo = 0, hi = m;    // leaves not heavier than the package
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (leaf_w(mid) <= w) lo = mid + 1;
                else hi = mid;
            }
            s_list[j + lo] = w;
        }
        src_len = m + np;
        __syncthreads();
    }
    // position of leaf i in the merged list of level l: i + #(packages lighter than it)
    if (t == 0) s_npk[kMaxCodeBits] = 0;
    __syncthreads();
    for (int k = t; k < kMaxCodeBits * m; k += nt) {
        const int l = 1 + k / m, i = k % m;
        const unsigned long long* pk = s_pk + (size_t)(l - 1) * kNumLitLen;
        const unsigned long long w = leaf_w(i);
        int a = 0, b = s_npk[l];
        while (a < b) {
            const int q = (a + b) >> 1;
            if (pk[q] < w) a = q + 1;
            else b = q;
        }
        s_pos[(l - 1) * kNumLitLen + i] = (uint16_t)(i + a);
    }
    __syncthreads();
    if (t == 0) {
        int sel = 2 * m - 2;
        for (int l = 1; l <= kMaxCodeBits; ++l) {
            // leaves among the first sel items (positions increase with i)
            const uint16_t* pos = s_pos + (l - 1) * kNumLitLen;
            int lo = 0, hi = m;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (pos[mid] < sel) lo = mid + 1;
                else hi = mid;
            }
            s_cnt[l] = lo;
            sel = 2 * (sel - lo);
        }
    }
    for (int i = t; i < kNumLitLen; i += nt) s_len[i] = 0;
    if (t < kMaxCodeBits + 2) s_blc[t] = 0;
    __syncthreads();
    if (t < m) {
        int L = 0;
        for (int l = 1; l <= kMaxCodeBits; ++l) L += t < s_cnt[l];
        s_len[(int)(s_key[t] & 0xFFFFu)] = (uint8_t)L;
        atomicAdd(&s_blc[L], 1);
    }
    __syncthreads();
    if (t == 0) {
        // canonical first codes per length
        int code = 0;
        s_blc[0] = 0;
        for (int L = 1; L <= kMaxCodeBits; ++L) {
            code = (code + s_blc[L - 1]) << 1;
            s_next[L] = code;
        }
    }
    __syncthreads();
    // a symbol's code: its length's first code plus its rank among the
    // symbols of that length (ballots within waves, counts across them)
    constexpr int kSymWaves = (kNumLitLen + 63) / 64;
    const int lane = t & 63, wave = t >> 6;
    const int myL = t < kNumLitLen ? s_len[t] : 0;
    int rank = 0;
    if (wave < kSymWaves) {
        for (int L = 1; L <= kMaxCodeBits; ++L) {
            const unsigned long long b = __ballot(myL == L);
            if (myL == L) rank = __popcll(b & ((1ull << lane) - 1ull));
            if (lane == 0) s_wcnt[wave][L] = __popcll(b);
        }
    }
    __syncthreads();
    if (t < kNumLitLen) {
        uint32_t v = 0;
        if (myL) {
            for (int q = 0; q < wave; ++q) rank += s_wcnt[q][myL];
            v = bit_reverse((uint32_t)(s_next[myL] + rank), myL) | ((uint32_t)myL << 24);
        }
        s_code[t] = v;
        ws.code[t] = v;
    }
    if (t < kHistStride / 2)
        s_len2[t] = (2 * t < kNumLitLen ? s_len[2 * t] : 0u) | ((2 * t + 1 < kNumLitLen ? s_len[2 * t + 1] : 0u) << 16);
    // 7. the block header, from bit 24 of the word that holds bytes 40-43
    if (t == 0) {
        s_hdr[0] = 'T' | (0x78u << 8) | (0x01u << 16);   // end of "IDAT", zlib header (deflate, 32K, fastest)
        auto put = [&](int pos, uint32_t bits, int len) {
            s_hdr[pos >> 5] |= bits << (pos & 31);
            if ((pos & 31) + len > 32) s_hdr[(pos >> 5) + 1] |= bits >> (32 - (pos & 31));
        };
        put(24, 5u, 3);                              // BFINAL = 1, BTYPE = 2 (dynamic)
        put(27, (uint32_t)(kNumLitLen - 257), 5);    // HLIT
        put(32, 0u, 5);                              // HDIST: one distance code
        put(37, 15u, 4);                             // HCLEN: all 19 code-length code lengths
        const int order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
        for (int q = 0; q < 19; ++q) put(41 + 3 * q, order[q] <= 15 ? 4u : 0u, 3);
        put(98 + 4 * kNumLitLen, bit_reverse(1u, 4), 4);   // the distance code: length 1
    }
    __syncthreads();
    if (t < kNumLitLen) {
        // code length l is sent as the 4-bit canonical code l (sixteen codes of length 4)
        const int pos = 98 + 4 * t;
        const uint32_t c = bit_reverse(s_len[t], 4);
        atomicOr(&s_hdr[pos >> 5], c << (pos & 31));
        if ((pos & 31) + 4 > 32) atomicOr(&s_hdr[(pos >> 5) + 1], c >> (32 - (pos & 31)));
    }
    __syncthreads();

    // 8. row bit lengths, Adler-32 partials and the scan
    const int per = (height + nt - 1) / nt;
    const int r0 = min(height, t * per), r1 = min(height, r0 + per);
    const uint64_t n = (uint64_t)stride + 1, N = n * (uint64_t)height;
    uint64_t sum = 0;
    uint32_t A = 0, B = 0;
    const uint32_t n_mod = (uint32_t)(n % kAdlerMod);
    for (int r = r0; r < r1; ++r) {
        uint32_t bits = ws.extra[r];
        const uint4* h = (const uint4*)(ws.hist + (size_t)r * kHistStride);
        for (int q = 0; q < kHistStride / 8; ++q) {
            const uint4 v = h[q];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t l2 = s_len2[q * 4 + k];
                bits += (w[k] & 0xFFFFu) * (l2 & 0xFFFFu) + (w[k] >> 16) * (l2 >> 16);
            }
        }
        ws.row_bits[r] = bits;
        sum += bits;
        // Adler-32 of the filtered rows: A = 1 + sum d, B = N + sum_i (N - i + 1) d_i;
        // row r starts at o = r n, and N - o - n = n (height - 1 - r)
        const uint32_t S = (uint32_t)ws.row_s[r];   // already reduced mod 65521
        const uint32_t coef = (uint32_t)(((uint64_t)n_mod * (uint32_t)((height - 1 - r) % kAdlerMod)) % kAdlerMod);
        A = (A + S) % kAdlerMod;
        B = (uint32_t)(((uint64_t)B + (uint64_t)coef * S + (uint32_t)ws.row_t[r]) % kAdlerMod);
    }
    s_part[t] = sum;
    s_a[t] = A;
    s_b[t] = B;
    __syncthreads();
    // inclusive scan (Hillis-Steele) of the bit counts; tree sums of A and B
    for (int o = 1; o < nt; o <<= 1) {
        const uint64_t y = t >= o ? s_part[t - o] : 0;
        uint64_t ya = 0, yb = 0;
        const bool red = (t & (2 * o - 1)) == 0 && t + o < nt;
        if (red) { ya = s_a[t + o]; yb = s_b[t + o]; }
        __syncthreads();
        s_part[t] += y;
        if (red) {
            s_a[t] = (uint32_t)(s_a[t] + ya) % kAdlerMod;
            s_b[t] = (uint32_t)(s_b[t] + yb) % kAdlerMod;
        }
        __syncthreads();
    }
    const uint64_t start = kBlockStartBit + kHeaderBits;
    uint64_t off = start + (t > 0 ? s_part[t - 1] : 0);
    uint32_t* words = (uint32_t*)png;
    for (int r = r0; r < r1; ++r) {
        const uint32_t nb = ws.row_bits[r];
        ws.row_off[r] = off;
        words[off >> 5] = 0u;
        words[(off + nb - 1) >> 5] = 0u;
        off += nb;
    }
    const uint64_t end = start + s_part[nt - 1];   // end of the row bits
    const int eob_len = (int)(s_code[256] >> 24);
    const uint64_t dbytes_end = (end + eob_len + 7) / 8;   // byte after the deflate data
    if (t == 0) {
        // the words of the end-of-block code and the padding after it
        for (uint64_t w = end >> 5; w <= (dbytes_end * 8 - 1) >> 5; ++w) words[w] = 0u;
    }
    // the zeroed words complete before any thread writes into them (one
    // workgroup: the barrier orders its global stores)
    __syncthreads();
    // bytes 40.. up to the first row: header words (the last one shares bits with row 0)
    const int hdr_words = (24 + kHeaderBits + 31) / 32;
    if (t < hdr_words) words[10 + t] = s_hdr[t];
    __syncthreads();
    if (t == 0) {
        const uint32_t eob = s_code[256] & 0xFFFFFFu;
        atomicOr(&words[end >> 5], eob << (end & 31));
        if ((end & 31) + eob_len > 32) atomicOr(&words[(end >> 5) + 1], eob >> (32 - (end & 31)));
    }
    __syncthreads();
    if (t == 0) {
        ws.totals[0] = end;
        const uint64_t Af = (1 + s_a[0]) % kAdlerMod, Bf = (N % kAdlerMod + s_b[0]) % kAdlerMod;
        const uint32_t adler = (uint32_t)((Bf << 16) | Af);
        ws.totals[3] = adler;
        uint8_t* a = png + dbytes_end;
        a[0] = (uint8_t)(adler >> 24); a[1] = (uint8_t)(adler >> 16); a[2] = (uint8_t)(adler >> 8);
        a[3] = (uint8_t)adler;
        const uint64_t zlib_len = dbytes_end + 4 - 41;
        png[33] = (uint8_t)(zlib_len >> 24); png[34] = (uint8_t)(zlib_len >> 16);
        png[35] = (uint8_t)(zlib_len >> 8); png[36] = (uint8_t)zlib_len;
        ws.totals[1] = zlib_len;
        ws.totals[4] = 0;   // CRC workgroup ticket
        // signature
        const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
        for (int i = 0; i < 8; ++i) png[i] = sig[i];
        // IHDR: length 13, type, width, height, depth 8, colour type 2 (RGB) / 6 (RGBA), 0, 0, 0, CRC
        uint8_t* h = png + 8;
        h[0] = 0; h[1] = 0; h[2] = 0; h[3] = 13;
        h[4] = 'I'; h[5] = 'H'; h[6] = 'D'; h[7] = 'R';
        h[8] = (uint8_t)(width >> 24); h[9] = (uint8_t)(width >> 16); h[10] = (uint8_t)(width >> 8); h[11] = (uint8_t)width;
        h[12] = (uint8_t)(height >> 24); h[13] = (uint8_t)(height >> 16); h[14] = (uint8_t)(height >> 8);
        h[15] = (uint8_t)height;
        h[16] = 8; h[17] = ch == 3 ? 2 : 6; h[18] = 0; h[19] = 0; h[20] = 0;
        h[21] = (uint8_t)(ihdr_crc >> 24); h[22] = (uint8_t)(ihdr_crc >> 16); h[23] = (uint8_t)(ihdr_crc >> 8);
        h[24] = (uint8_t)ihdr_crc;
        // IDAT type (its last byte, 'T', went out with the header words)
        png[37] = 'I'; png[38] = 'D'; png[39] = 'A';
    }
}

// Emit: one workgroup per row stages the filtered row and the code table in
// LDS, builds the row's bits there, then writes them at the row's bit offset:
// interior words whole, the first and last words (shared with the
// neighbouring rows or the bytes around the row data, and zeroed by the scan)
// through atomic OR.
__global__ void __launch_bounds__(kRowThreads) k_png_emit(int height, int stride, PngWs ws, uint8_t* png) {
    extern __shared__ uint32_t s_mem[];
    __shared__ uint32_t s_u[kRowWaves];
    __shared__ int s_i[kRowWaves];
    __shared__ uint32_t s_code[kNumLitLen];
    const int r = blockIdx.x;
    if (r >= height) return;
    const int n = stride + 1;
    const uint32_t row_words = ((uint32_t)n + 3) / 4;
    uint8_t* s_row = (uint8_t*)s_mem;
    uint32_t* s_bits = s_mem + row_words;
    const uint8_t* row = ws.filt + (size_t)r * n;
    const uint32_t nbits = ws.row_bits[r];
    const uint32_t nwords = (nbits + 31) / 32 + 1;
    for (int i = threadIdx.x; i < n; i += kRowThreads) s_row[i] = row[i];
    for (uint32_t i = threadIdx.x; i < nwords; i += kRowThreads) s_bits[i] = 0u;
    for (int i = threadIdx.x; i < kNumLitLen; i += kRowThreads) s_code[i] = ws.code[i];
    __syncthreads();
    auto put = [&](uint32_t pos, uint32_t bits, int len) {
        // len <= 32 (a match: at most 12 + 5 + 1 bits)
        const uint32_t w = pos >> 5, o = pos & 31;
        atomicOr(&s_bits[w], bits << o);
        if (o + len > 32) atomicOr(&s_bits[w + 1], bits >> (32 - o));
    };
    // bits of a run: the same symbols the histogram counted
    auto run_len = [&](int v, int len) {
        uint32_t b = 0;
        run_symbols(
            v, len, [&](int c, int count) { b += (s_code[c] >> 24) * (uint32_t)count; },
            [&](int L) {
                int sym, ne, base;
                len_symbol(L, sym, ne, base);
                b += (s_code[sym] >> 24) + (uint32_t)ne + 1u;
            });
        return b;
    };
    const RowRuns R = row_runs(s_row, n, s_i);
    uint32_t mine = 0;
    visit_runs(s_row, R, [&](int v, int len) { mine += run_len(v, len); });
    uint32_t total;
    uint32_t pos = block_excl_sum(mine, s_u, total);
    visit_runs(s_row, R, [&](int v, int len) {
        run_symbols(
            v, len,
            [&](int c, int count) {
                const uint32_t code = s_code[c];
                const int l = (int)(code >> 24);
                for (int k = 0; k < count; ++k) {
                    put(pos, code & 0xFFFFFFu, l);
                    pos += l;
                }
            },
            [&](int L) {
                int sym, ne, base;
                len_symbol(L, sym, ne, base);
                const uint32_t code = s_code[sym];
                const int l = (int)(code >> 24);
                // length code, extra bits, then the distance code: one 0 bit
                put(pos, (code & 0xFFFFFFu) | ((uint32_t)(L - base) << l), l + ne);
                pos += l + ne + 1;
            });
    });
    __syncthreads();
    const uint64_t off = ws.row_off[r];
    uint32_t* out = (uint32_t*)png;
    const uint64_t w_first = off >> 5, w_last = (off + nbits - 1) >> 5;
    const uint32_t sh = (uint32_t)(off & 31);
    for (uint64_t w = w_first + threadIdx.x; w <= w_last; w += kRowThreads) {
        const uint64_t k = w - w_first;   // output word k of this row
        // bits [32k - sh, 32k - sh + 32) of the row stream
        uint32_t val;
        if (sh == 0) val = s_bits[k];
        else {
            const uint32_t lo = k > 0 ? s_bits[k - 1] >> (32 - sh) : 0u;
            const uint32_t hi = k < nwords ? s_bits[k] << sh : 0u;
            val = lo | hi;
        }
        if (w == w_first || w == w_last) atomicOr(&out[w], val);
        else out[w] = val;
    }
}

// CRC-32 of the IDAT type + data (len bytes from png + 37), in workgroups of
// kCrcBlock bytes. The stream is taken as nblk * kCrcBlock bytes with
// pad = nblk * kCrcBlock - len zero bytes in front: a CRC register at zero
// stays zero over zero bytes, and the standard initial value 0xFFFFFFFF is
// XORed into the register at the first real byte, so the result is the CRC
// of the real bytes. Every workgroup and every thread then covers a
// power-of-two span, and spans combine as crc(A B) = crc(A) x^(8|B|) + crc(B)
// with the powers taken from the table. The last workgroup to finish (atomic
// ticket; nothing waits) combines the workgroups' CRCs and writes the CRC,
// IEND and the PNG length.
__global__ void __launch_bounds__(256) k_crc(uint8_t* png, PngWs ws, uint64_t max_blocks, uint64_t* png_len,
                                             PowTable pw) {
    __shared__ uint32_t table[256];
    __shared__ uint32_t s_w[4];
    __shared__ int s_last;
    constexpr int kPer = kCrcBlock / 256;   // 16 bytes per thread
    static_assert(kPer == 16 && kCrcBlock == 4096, "the tree levels below assume 16-byte spans in 4 KiB blocks");
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        uint32_t c = (uint32_t)i;
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? (c >> 1) ^ kCrcPoly : c >> 1;
        table[i] = c;
    }
    __syncthreads();
    const uint64_t len = 4 + ws.totals[1];
    const uint64_t nblk = (len + kCrcBlock - 1) / kCrcBlock;
    const uint64_t blk = blockIdx.x;
    if (blk >= max_blocks || blk >= nblk) return;
    const uint64_t pad = nblk * kCrcBlock - len;
    const uint8_t* base = png + 37;
    const uint64_t q0 = blk * kCrcBlock + (uint64_t)threadIdx.x * kPer;
    uint32_t c = 0;
    for (int k = 0; k < kPer; ++k) {
        const uint64_t q = q0 + k;
        if (q == pad) c ^= 0xFFFFFFFFu;
        const uint8_t b = q >= pad ? base[q - pad] : (uint8_t)0;
        c = crc_update_raw(c, b, table);
    }
    // tree over the wave: level j joins spans of 16 * 2^j bytes
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < 6; ++j) {
        const uint32_t right = __shfl_down(c, 1 << j, 64);
        c = gf2_mulmod(c, pw.p[4 + j]) ^ right;   // meaningful on lanes that are multiples of 2^(j+1)
    }
    if (lane == 0) s_w[wave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = s_w[0];
        for (int w = 1; w < 4; ++w) b = gf2_mulmod(b, pw.p[10]) ^ s_w[w];   // 1024-byte spans
        ws.crc_part[blk] = b;
        __threadfence();
        const unsigned long long tk = atomicAdd((unsigned long long*)&ws.totals[4], 1ull);
        s_last = tk == nblk - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    // the last workgroup: thread t folds blocks [t m, (t + 1) m) by Horner's
    // rule, moves the result past the blocks after them, then XOR over threads
    const uint64_t m = (nblk + 255) / 256;
    const uint64_t b0 = min(nblk, (uint64_t)threadIdx.x * m), b1 = min(nblk, b0 + m);
    uint32_t acc = 0;
    for (uint64_t b = b0; b < b1; ++b)
        acc = gf2_mulmod(acc, pw.p[12]) ^ __hip_atomic_load(&ws.crc_part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t v = b1 > b0 ? gf2_mulmod(acc, x_pow_8n_tab(pw, (nblk - b1) * (uint64_t)kCrcBlock)) : 0u;
    for (int o = 32; o > 0; o >>= 1) v ^= __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) s_w[wave] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const uint32_t crc = (s_w[0] ^ s_w[1] ^ s_w[2] ^ s_w[3]) ^ 0xFFFFFFFFu;   // final XOR
    uint8_t* cp = png + 37 + len;
    cp[0] = (uint8_t)(crc >> 24); cp[1] = (uint8_t)(crc >> 16); cp[2] = (uint8_t)(crc >> 8); cp[3] = (uint8_t)crc;
    const uint8_t iend[12] = {0, 0, 0, 0, 'I', 'E', 'N', 'D', 0xAE, 0x42, 0x60, 0x82};
    for (int i = 0; i < 12; ++i) cp[4 + i] = iend[i];
    const uint64_t total = 37 + len + 4 + 12;
    ws.totals[2] = total;
    *png_len = total;
}

__global__ void __launch_bounds__(256) k_base64(const uint8_t* __restrict__ in, const uint64_t* in_len, uint64_t in_cap,
                                                uint8_t* __restrict__ out, uint64_t* out_len) {
    const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    uint64_t n = *in_len;
    if (n > in_cap) n = in_cap;
    const uint64_t groups = (n + 2) / 3;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = 3 * g;
        const uint32_t b0 = in[i];
        const uint32_t b1 = i + 1 < n ? in[i + 1] : 0u;
        const uint32_t b2 = i + 2 < n ? in[i + 2] : 0u;
        const uint32_t v = (b0 << 16) | (b1 << 8) | b2;
        uint8_t* o = out + 4 * g;
        o[0] = (uint8_t)tab[(v >> 18) & 63];
        o[1] = (uint8_t)tab[(v >> 12) & 63];
        o[2] = i + 1 < n ? (uint8_t)tab[(v >> 6) & 63] : (uint8_t)'=';
        o[3] = i + 2 < n ? (uint8_t)tab[v & 63] : (uint8_t)'=';
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_len = 4 * groups;
}

// host-side CRC of the IHDR chunk (type + 13 data bytes)
uint32_t host_crc32(const uint8_t* p, size_t n) {
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? (c >> 1) ^ kCrcPoly : c >> 1;
    }
    return c ^ 0xFFFFFFFFu;
}

struct WsLayout {
    size_t filt, row_bits, row_s, row_t, row_off, totals, crc_part, hist, extra, code, bytes;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

uint64_t png_bound(int w, int h, int ch) {
    // no byte costs more than kMaxCodeBits (a literal; a match of 3 or more
    // bytes costs at most kMaxCodeBits + 6)
    const uint64_t n = (uint64_t)w * ch + 1;
    const uint64_t bits = (uint64_t)kHeaderBits + (uint64_t)h * n * kMaxCodeBits + kMaxCodeBits + 7;
    return kHeadBytes + 2 + (bits + 7) / 8 + 4 + 4 + 12 + 8;   // + slack for the last shared word
}

WsLayout ws_layout(int w, int h, int ch, uint64_t bound) {
    WsLayout L;
    size_t o = 0;
    L.filt = o; o = align256(o + (size_t)h * ((size_t)w * ch + 1));
    L.row_bits = o; o = align256(o + (size_t)h * 4);
    L.row_s = o; o = align256(o + (size_t)h * 8);
    L.row_t = o; o = align256(o + (size_t)h * 8);
    L.row_off = o; o = align256(o + (size_t)h * 8);
    L.totals = o; o = align256(o + 8 * 8);
    L.crc_part = o; o = align256(o + (size_t)((bound + kCrcBlock - 1) / kCrcBlock + 1) * 4);
    L.hist = o; o = align256(o + (size_t)h * kHistStride * 2);
    L.extra = o; o = align256(o + (size_t)h * 4);
    L.code = o; o = align256(o + (size_t)kNumLitLen * 4);
    L.bytes = o;
    return L;
}

bool bad_dims(int w, int h, int ch) {
    return w <= 0 || h <= 0 || (ch != 3 && ch != 4) || (int64_t)w * ch > kMaxRowBytes || h > 65536;
}

}  // namespace

extern "C" uint64_t mk_png_max_bytes(int width, int height, int channels) {
    if (bad_dims(width, height, channels)) return 0;
    return png_bound(width, height, channels);
}

extern "C" uint64_t mk_png_workspace_bytes(int width, int height, int channels) {
    if (bad_dims(width, height, channels)) return 0;
    return ws_layout(width, height, channels, png_bound(width, height, channels)).bytes;
}

extern "C" mk_status mk_png_encode_async(const uint8_t* dev_pixels, int width, int height, int channels,
                                         void* dev_workspace, uint64_t workspace_bytes, uint8_t* dev_png,
                                         uint64_t png_capacity, uint64_t* dev_png_len, void* stream) {
    if (bad_dims(width, height, channels))
        return mk_fail(nullptr, MK_ERR_ARG, "mk_png_encode_async: bad image %dx%dx%d (channels 3 or 4, row <= %d bytes)",
                       width, height, channels, kMaxRowBytes);
    if (!dev_pixels || !dev_workspace || !dev_png || !dev_png_len)
        return mk_fail(nullptr, MK_ERR_ARG, "mk_png_encode_async: null buffer");
    const uint64_t bound = png_bound(width, height, channels);
    if (png_capacity < bound)
        return mk_fail(nullptr, MK_ERR_ARG, "mk_png_encode_async: capacity %llu below the bound %llu",
                       (unsigned long long)png_capacity, (unsigned long long)bound);
    if (((uintptr_t)dev_png & 3) != 0) return mk_fail(nullptr, MK_ERR_ARG, "mk_png_encode_async: output not 4-byte aligned");
    const WsLayout L = ws_layout(width, height, channels, bound);
    if (workspace_bytes < L.bytes || ((uintptr_t)dev_workspace & 255) != 0)
        return mk_fail(nullptr, MK_ERR_ARG, "mk_png_encode_async: workspace needs %zu bytes, 256-byte aligned", L.bytes);
    uint8_t* w = (uint8_t*)dev_workspace;
    PngWs ws;
    ws.filt = w + L.filt;
    ws.row_bits = (uint32_t*)(w + L.row_bits);
    ws.row_s = (uint64_t*)(w + L.row_s);
    ws.row_t = (uint64_t*)(w + L.row_t);
    ws.row_off = (uint64_t*)(w + L.row_off);
    ws.totals = (uint64_t*)(w + L.totals);
    ws.crc_part = (uint32_t*)(w + L.crc_part);
    ws.hist = (uint16_t*)(w + L.hist);
    ws.extra = (uint32_t*)(w + L.extra);
    ws.code = (uint32_t*)(w + L.code);
    // IHDR CRC on the host (the header depends only on the dims)
    uint8_t ih[17] = {'I', 'H', 'D', 'R', (uint8_t)(width >> 24), (uint8_t)(width >> 16), (uint8_t)(width >> 8),
                      (uint8_t)width, (uint8_t)(height >> 24), (uint8_t)(height >> 16), (uint8_t)(height >> 8),
                      (uint8_t)height, 8, (uint8_t)(channels == 3 ? 2 : 6), 0, 0, 0};
    const uint32_t ihdr_crc = host_crc32(ih, sizeof(ih));
    const int stride = width * channels;
    hipStream_t s = (hipStream_t)stream;
    const size_t row_lds = ((size_t)stride + 1 + 3) / 4 * 4;
    hipLaunchKernelGGL(k_png_filter_hist, dim3(height), dim3(kRowThreads), row_lds, s, dev_pixels, width, height,
                       channels, ws);
    hipLaunchKernelGGL(k_huff_scan, dim3(1), dim3(kScanThreads), 0, s, height, stride, ws, dev_png, ihdr_crc, width,
                       channels);
    const size_t emit_lds = row_lds + ((size_t)(stride + 1) * kMaxCodeBits / 32 + 2) * 4;
    hipLaunchKernelGGL(k_png_emit, dim3(height), dim3(kRowThreads), emit_lds, s, height, stride, ws, dev_png);
    const uint64_t max_blocks = (bound + kCrcBlock - 1) / kCrcBlock;
    static const PowTable pw = make_pow_table();
    hipLaunchKernelGGL(k_crc, dim3((uint32_t)max_blocks), dim3(256), 0, s, dev_png, ws, max_blocks, dev_png_len, pw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mk_hip_fail(nullptr, e, "mk_png_encode_async launch");
    return MK_OK;
}

extern "C" uint64_t mk_base64_max_bytes(uint64_t n) { return 4 * ((n + 2) / 3); }

extern "C" mk_status mk_base64_encode_async(const uint8_t* dev_in, const uint64_t* dev_in_len, uint64_t in_capacity,
                                            uint8_t* dev_out, uint64_t out_capacity, uint64_t* dev_out_len,
                                            void* stream) {
    if (!dev_in || !dev_in_len || !dev_out || !dev_out_len)
        return mk_fail(nullptr, MK_ERR_ARG, "mk_base64_encode_async: null buffer");
    if (out_capacity < mk_base64_max_bytes(in_capacity))
        return mk_fail(nullptr, MK_ERR_ARG, "mk_base64_encode_async: output capacity below 4 * ceil(in_capacity / 3)");
    const uint64_t groups = (in_capacity + 2) / 3;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((groups + 255) / 256 + 1, 65535);
    hipLaunchKernelGGL(k_base64, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dev_in, dev_in_len, in_capacity,
                       dev_out, dev_out_len);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mk_hip_fail(nullptr, e, "mk_base64_encode_async launch");
    return MK_OK;
}
