// mk_lattice.h -- the ray casters' walk over empty cells: the result of the
// float accumulation acc += l1, repeated while the next point lies below a
// bound, taken in a few steps instead of one per point. Plain C++ apart from
// the __host__ __device__ qualifiers, so that a host test can compile it and
// compare it with the loop it replaces.
#pragma once

#include <stdint.h>

namespace mkr {

// Advances acc and it exactly as
//   while (acc + l1 < a_skip && it < max_iter) { acc += l1; ++it; }
// does (rl1 = 1 / l1 serves only to estimate). In the binade [2^e, 2^(e+1))
// the floats are the multiples of u = 2^(e-23), and acc is one of them, so
// while a sum stays at or below 2^(e+1) - u, fl(acc + l1) = acc + D*u with
// D = q rounded to the nearest integer, q = l1/u: n additions land exactly on
// acc + n*D*u. One jump takes n of them and the loop itself takes the rest
// one at a time, so it ends where the loop ends. n is the float estimate of
// the count, capped at min(q/4, 2^20), less one when its last addition would
// reach a_skip. The count is ceil(x) - 1 for x = (a_skip/u - a)/D, and the
// estimate is within x * (1/(2q) + 2^-22) of x: under the cap within 0.44, so
// never more than one too high. No jump is made when it would leave the
// binade (the loop crosses into the next one), when q lies halfway between
// two integers (the rounding would then depend on acc's last bit), when q is
// below 8 or not below 2^24, or when acc is negative or below 2^-100.
__host__ __device__ inline void lattice_skip(float& acc, float l1, float rl1, float a_skip, int& it, int max_iter) {
    const float nf = (a_skip - acc) * rl1;   // about the additions below a_skip
    if (nf >= 2.0f) {
        const uint32_t bits = __builtin_bit_cast(uint32_t, acc);
        const int be = (int)(bits >> 23);   // biased exponent; >= 256 when acc is negative
        const int sh = 150 - be;            // u = 2^-sh
        const float q = __builtin_ldexpf(l1, sh);
        const bool ok = be >= 27 && be <= 254 && q >= 8.0f && q < 16777216.0f && q - __builtin_floorf(q) != 0.5f;
        const int D = ok ? (int)__builtin_rintf(q) : 0;
        const int a = (int)((bits & 0x7FFFFFu) | 0x800000u);   // acc / u
        const float cap = q < 4194304.0f ? 0.25f * q : 1048576.0f;
        float ne = nf < cap ? nf : cap;
        const float left = (float)(max_iter - it);
        ne = ne < left ? ne : left;
        int n = ok ? (int)ne : 0;
        // the float product n*D is exact below 2^24 and at least 2^24 above;
        // a jump that would leave the binade is not made
        const float P = (float)n * (float)D;
        const float S = __builtin_ldexpf(a_skip, sh);   // a_skip / u
        if ((float)a + P >= S) --n;
        if (P > (float)(16777215 - a)) n = 0;
        if (n > 0) {
            const uint32_t m = (uint32_t)(a + n * D);
            acc = __builtin_bit_cast(float, (bits & 0xFF800000u) | (m & 0x7FFFFFu));
            it += n;
        }
    }
    while (acc + l1 < a_skip && it < max_iter) {
        acc += l1;
        ++it;
    }
}

}  // namespace mkr
