// Unsigned integers of W 32-bit limbs (least significant first), for the exact parts of
// the operations the reference runs on the host in MPFR: the division and square root
// of multiple-precision numbers (src/arith/div.cuh, src/arith/sqrt.cuh). These run one
// thread per number and are not on any hot path; every loop is bounded by W.
//
// __host__ __device__ so the same code is checked on the CPU against Python integers.
#pragma once

#include <stdint.h>

#include "numerics.hpp"

namespace mph {

template <int W>
struct Big {
    uint32_t l[W];
};

template <int W>
MPH_HD void big_zero(Big<W> &a) {
#pragma unroll
    for (int i = 0; i < W; i++) a.l[i] = 0;
}
template <int W>
MPH_HD bool big_is_zero(const Big<W> &a) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < W; i++) o |= a.l[i];
    return o == 0;
}
MPH_HD int clz32(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clz((int)v);
#else
    return v ? __builtin_clz(v) : 32;
#endif
}
MPH_HD int ctz32(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return v ? __ffs((int)v) - 1 : 32;
#else
    return v ? __builtin_ctz(v) : 32;
#endif
}
// Every index below is a compile-time constant (loops over limbs are unrolled, dynamic
// positions are resolved with selects), so a Big stays in registers: an array indexed
// with a run-time value would be placed in scratch memory.

// number of significant bits (0 for zero)
template <int W>
MPH_HD int big_bitlen(const Big<W> &a) {
    int r = 0;
#pragma unroll
    for (int i = 0; i < W; i++) r = a.l[i] ? 32 * i + 32 - clz32(a.l[i]) : r;
    return r;
}
// number of trailing zero bits (0 for zero)
template <int W>
MPH_HD int big_ctz(const Big<W> &a) {
    int r = 0;
#pragma unroll
    for (int i = W - 1; i >= 0; i--) r = a.l[i] ? 32 * i + ctz32(a.l[i]) : r;
    return r;
}
template <int W>
MPH_HD int big_bit(const Big<W> &a, int i) {
    const int q = i >> 5;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < W; k++) v = q == k ? a.l[k] : v;
    return (i >= 0 && i < 32 * W) ? (int)((v >> (i & 31)) & 1u) : 0;
}
// any nonzero bit below position i
template <int W>
MPH_HD bool big_any_below(const Big<W> &a, int i) {
    uint32_t o = 0;
#pragma unroll
    for (int q = 0; q < W; q++) {
        const int k = i - 32 * q;  // bits of limb q below position i
        const uint32_t mask = k >= 32 ? 0xffffffffu : (k <= 0 ? 0u : (1u << (k & 31)) - 1u);
        o |= a.l[q] & mask;
    }
    return o != 0;
}
// a << s for s >= 0, bits shifted beyond W limbs are lost: the whole-limb part as a
// barrel of conditional moves (1, 2, 4, ... limbs), then the bit part
template <int W>
MPH_HD Big<W> big_shl(const Big<W> &a, int s) {
    Big<W> r = a;
    const int q = s >> 5, b = s & 31;
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
        const bool on = (q & k) != 0;
#pragma unroll
        for (int i = W - 1; i >= 0; i--) r.l[i] = on ? (i >= k ? r.l[i - k] : 0u) : r.l[i];
    }
    const bool out = q >= W;
#pragma unroll
    for (int i = W - 1; i >= 0; i--) {
        const uint32_t lo = i > 0 ? r.l[i - 1] : 0u;
        r.l[i] = out ? 0u : (b ? (r.l[i] << b) | (lo >> ((32 - b) & 31)) : r.l[i]);
    }
    return r;
}
// a >> s for s >= 0
template <int W>
MPH_HD Big<W> big_shr(const Big<W> &a, int s) {
    Big<W> r = a;
    const int q = s >> 5, b = s & 31;
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
        const bool on = (q & k) != 0;
#pragma unroll
        for (int i = 0; i < W; i++) r.l[i] = on ? (i + k < W ? r.l[i + k] : 0u) : r.l[i];
    }
    const bool out = q >= W;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t hi = i + 1 < W ? r.l[i + 1] : 0u;
        r.l[i] = out ? 0u : (b ? (r.l[i] >> b) | (hi << ((32 - b) & 31)) : r.l[i]);
    }
    return r;
}
template <int W>
MPH_HD int big_cmp(const Big<W> &a, const Big<W> &b) {
    int r = 0;
#pragma unroll
    for (int i = 0; i < W; i++) r = a.l[i] != b.l[i] ? (a.l[i] > b.l[i] ? 1 : -1) : r;
    return r;
}
// a -= b (a >= b)
template <int W>
MPH_HD void big_sub(Big<W> &a, const Big<W> &b) {
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint64_t t = (uint64_t)a.l[i] - b.l[i] - borrow;
        a.l[i] = (uint32_t)t;
        borrow = (t >> 63) & 1u;
    }
}
// a += b
template <int W>
MPH_HD void big_add(Big<W> &a, const Big<W> &b) {
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint64_t t = (uint64_t)a.l[i] + b.l[i] + carry;
        a.l[i] = (uint32_t)t;
        carry = t >> 32;
    }
}
template <int W>
MPH_HD void big_add_u32(Big<W> &a, uint32_t v) {
    uint64_t carry = v;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint64_t t = (uint64_t)a.l[i] + carry;
        a.l[i] = (uint32_t)t;
        carry = t >> 32;
    }
}
// a = a * m + c
template <int W>
MPH_HD void big_mul_add_u32(Big<W> &a, uint32_t m, uint32_t c) {
    uint64_t carry = c;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint64_t t = (uint64_t)a.l[i] * m + carry;
        a.l[i] = (uint32_t)t;
        carry = t >> 32;
    }
}
// a * b, exactly
template <int A, int B>
MPH_HD Big<A + B> big_mul(const Big<A> &a, const Big<B> &b) {
    Big<A + B> r;
    big_zero(r);
#pragma unroll
    for (int i = 0; i < A; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < B; j++) {
            const uint64_t t = (uint64_t)a.l[i] * b.l[j] + r.l[i + j] + c;
            r.l[i + j] = (uint32_t)t;
            c = t >> 32;
        }
        r.l[i + B] = (uint32_t)c;
    }
    return r;
}
// a in a wider (or equal) type
template <int WB, int WA>
MPH_HD Big<WB> big_widen(const Big<WA> &a) {
    static_assert(WB >= WA, "widen");
    Big<WB> r;
#pragma unroll
    for (int i = 0; i < WB; i++) r.l[i] = i < WA ? a.l[i] : 0u;
    return r;
}
// a mod m for a modulus below 2^31
template <int W>
MPH_HD uint32_t big_mod_u32(const Big<W> &a, uint32_t m) {
    uint64_t r = 0;
#pragma unroll
    for (int i = W - 1; i >= 0; i--) r = ((r << 32) | a.l[i]) % m;
    return (uint32_t)r;
}
// a mod m through exact FP64 steps (modd), 16 bits at a time: r 2^16 + digit < 2^47
template <int W>
MPH_HD int big_mod_d(const Big<W> &a, double m, double minv) {
    int r = 0;
#pragma unroll
    for (int i = W - 1; i >= 0; i--) {
        r = modd((double)r * 65536.0 + (double)(a.l[i] >> 16), m, minv);
        r = modd((double)r * 65536.0 + (double)(a.l[i] & 0xffffu), m, minv);
    }
    return r;
}
// a rounded to p significant bits, to nearest with ties to even (the value a * 2^e is
// kept: e grows by the bits dropped); the result may be 2^p
template <int W>
MPH_HD void big_round(Big<W> &a, int &e, int p) {
    const int b = big_bitlen(a);
    if (b <= p) return;
    const int sh = b - p;
    const int half = big_bit(a, sh - 1), sticky = big_any_below(a, sh - 1) ? 1 : 0;
    a = big_shr(a, sh);
    if (half && (sticky || (a.l[0] & 1u))) big_add_u32(a, 1u);
    e += sh;
}
// floor(a / b * 2^(q - 1)) for a, b of equal bit length L (so 1/2 < a/b < 2), with
// W >= L + 2 bits: q restoring-division steps; returns the quotient and whether the
// division left a remainder
template <int W>
MPH_HD Big<W> big_div_bits(const Big<W> &a, const Big<W> &b, int q, bool &inexact) {
    Big<W> r = a, quo;
    big_zero(quo);
    for (int i = 0; i < q; i++) {
        quo = big_shl(quo, 1);
        if (big_cmp(r, b) >= 0) {
            big_sub(r, b);
            quo.l[0] |= 1u;
        }
        r = big_shl(r, 1);
    }
    inexact = !big_is_zero(r);
    return quo;
}
// s = floor(sqrt(a)) by the digit-by-digit method; inexact when s^2 != a
template <int W>
MPH_HD Big<W> big_isqrt(const Big<W> &a, bool &inexact) {
    Big<W> num = a, res, bit, t;
    big_zero(res);
    big_zero(bit);
    const int b = big_bitlen(a);
    if (b == 0) {
        inexact = false;
        return res;
    }
    const int top = (b - 1) & ~1;  // highest even bit position not above a's leading bit
    bit.l[top >> 5] = 1u << (top & 31);
    for (int i = top; i >= 0; i -= 2) {
        t = res;
        big_add(t, bit);
        res = big_shr(res, 1);
        if (big_cmp(num, t) >= 0) {
            big_sub(num, t);
            big_add(res, bit);
        }
        bit = big_shr(bit, 2);
    }
    inexact = !big_is_zero(num);
    return res;
}

}  // namespace mph
