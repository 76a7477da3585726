// BN254 Fq / Fr arithmetic for gfx950, usable on the host as well.
//
// Register form: nine limbs of 29 bits (radix 2^29), Montgomery radix
// R = 2^261.  A Montgomery product accumulates every column of 29x29-bit
// partial products in one 64-bit register with v_mad_u64_u32 and never needs
// a carry flag: a column holds at most 18 products below 2^58 (2^60 when one
// operand is "wide", see below), so it cannot overflow.  Measured on MI355X
// this runs 1.75x faster than 8x32-bit CIOS, whose carry chains cost three to
// four extra VALU instructions per partial product.
//
// Memory form: eight little-endian 32-bit words (256 bits).  Any value below
// 2^256 round-trips, so lazily reduced values (< 2m) can be stored as is.
//
// Invariants used throughout:
//   * "normalized": every limb < 2^29.
//   * mul/sqr outputs are normalized and < 2m.
//   * mul(a, b) accepts one "wide" operand a whose limbs are < 2^31 (the sum
//     or SUB4-difference of normalized values) and a normalized b, provided
//     a*b < m * 2^261; then the output is < 2m.
//   * reduce(x) maps any normalized x < 2^261 to a normalized value < 2m.
#pragma once
#include <stdint.h>
#include "bn254_constants.hpp"

#if defined(__HIPCC__)
#define HD __host__ __device__ __forceinline__
#else
#define HD inline
#endif

namespace bn254 {

template <class C>
struct Fp {
    uint32_t l[LIMBS];
};

using Fq = Fp<FqParams>;
using Fr = Fp<FrParams>;

template <class C>
HD Fp<C> fp_const(const uint32_t (&v)[LIMBS]) {
    Fp<C> r;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) r.l[i] = v[i];
    return r;
}

template <class C>
HD Fp<C> fp_zero() {
    Fp<C> r;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) r.l[i] = 0;
    return r;
}

template <class C>
HD Fp<C> fp_one() { return fp_const<C>(C::ONE); }

// Low 32 bits of a*b.  On the device through v_mad_u64_u32, which issues at
// full rate (the 32-bit v_mul_lo_u32 is quarter rate): the Montgomery digit
// of every column is on the serial chain of a product.
HD uint32_t mul_lo32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r, carry;
    asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(carry) : "v"(a), "v"(b));
    return uint32_t(r);
#else
    return a * b;
#endif
}

// Montgomery product a*b/R mod m, FIPS (product scanning) order.
template <class C>
HD Fp<C> mul(const Fp<C>& a, const Fp<C>& b) {
    uint32_t m[LIMBS];
    Fp<C> out;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
            acc += uint64_t(a.l[i]) * b.l[k - i];
            acc += uint64_t(m[i]) * C::MOD[k - i];
        }
        acc += uint64_t(a.l[k]) * b.l[0];
        m[k] = mul_lo32(uint32_t(acc), C::INV) & LIMB_MASK;
        acc += uint64_t(m[k]) * C::MOD[0];
        acc >>= LIMB_BITS;
    }
#pragma unroll
    for (int k = LIMBS; k < 2 * LIMBS - 1; ++k) {
#pragma unroll
        for (int i = k - LIMBS + 1; i < LIMBS; ++i) {
            acc += uint64_t(a.l[i]) * b.l[k - i];
            acc += uint64_t(m[i]) * C::MOD[k - i];
        }
        out.l[k - LIMBS] = uint32_t(acc) & LIMB_MASK;
        acc >>= LIMB_BITS;
    }
    out.l[LIMBS - 1] = uint32_t(acc);
    return out;
}

// Montgomery square of a normalized a: off-diagonal products use 2*a_i.
template <class C>
HD Fp<C> sqr(const Fp<C>& a) {
    uint32_t d[LIMBS];
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) d[i] = a.l[i] << 1;
    uint32_t m[LIMBS];
    Fp<C> out;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 2 * LIMBS - 1; ++k) {
#pragma unroll
        for (int i = (k < LIMBS ? 0 : k - LIMBS + 1); i < (k + 1) / 2; ++i)
            acc += uint64_t(d[i]) * a.l[k - i];
        if ((k & 1) == 0) acc += uint64_t(a.l[k / 2]) * a.l[k / 2];
#pragma unroll
        for (int i = (k < LIMBS ? 0 : k - LIMBS + 1); i < (k < LIMBS ? k : LIMBS); ++i)
            acc += uint64_t(m[i]) * C::MOD[k - i];
        if (k < LIMBS) {
            m[k] = mul_lo32(uint32_t(acc), C::INV) & LIMB_MASK;
            acc += uint64_t(m[k]) * C::MOD[0];
        } else {
            out.l[k - LIMBS] = uint32_t(acc) & LIMB_MASK;
        }
        acc >>= LIMB_BITS;
    }
    out.l[LIMBS - 1] = uint32_t(acc);
    return out;
}

// (a*b + c*d)/R mod m with one Montgomery reduction: a may be wide (limbs
// < 2^31), b, c, d normalized, and a*b + c*d < 44 m^2 (then the output is
// normalized and < 2m).  A column then holds at most 9 products below 2^60
// and 18 below 2^58 (< 2^63.8): no overflow.
template <class C>
HD Fp<C> mul2(const Fp<C>& a, const Fp<C>& b, const Fp<C>& c, const Fp<C>& d) {
    uint32_t m[LIMBS];
    Fp<C> out;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
            acc += uint64_t(a.l[i]) * b.l[k - i];
            acc += uint64_t(c.l[i]) * d.l[k - i];
            acc += uint64_t(m[i]) * C::MOD[k - i];
        }
        acc += uint64_t(a.l[k]) * b.l[0];
        acc += uint64_t(c.l[k]) * d.l[0];
        m[k] = mul_lo32(uint32_t(acc), C::INV) & LIMB_MASK;
        acc += uint64_t(m[k]) * C::MOD[0];
        acc >>= LIMB_BITS;
    }
#pragma unroll
    for (int k = LIMBS; k < 2 * LIMBS - 1; ++k) {
#pragma unroll
        for (int i = k - LIMBS + 1; i < LIMBS; ++i) {
            acc += uint64_t(a.l[i]) * b.l[k - i];
            acc += uint64_t(c.l[i]) * d.l[k - i];
            acc += uint64_t(m[i]) * C::MOD[k - i];
        }
        out.l[k - LIMBS] = uint32_t(acc) & LIMB_MASK;
        acc >>= LIMB_BITS;
    }
    out.l[LIMBS - 1] = uint32_t(acc);
    return out;
}

// Carry-propagate non-negative limbs (each < 2^32) into normalized form.
template <class C>
HD Fp<C> norm(Fp<C> x) {
#pragma unroll
    for (int i = 0; i < LIMBS - 1; ++i) {
        x.l[i + 1] += x.l[i] >> LIMB_BITS;
        x.l[i] &= LIMB_MASK;
    }
    return x;
}

// Normalized x < 2^261  ->  normalized value < 2m, congruent mod m.
// q = floor(top / ((m >> 232) + 1)) never exceeds x / m, and x - q*m < 2m.
// x - q*m is computed as x + q*(2^261 - m) mod 2^261, so no borrows occur.
template <class C>
HD Fp<C> reduce(const Fp<C>& x) {
    const uint32_t q = x.l[LIMBS - 1] / C::TOPDIV;
    Fp<C> r;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) {
        acc += uint64_t(q) * C::NEG[i] + x.l[i];
        r.l[i] = uint32_t(acc) & LIMB_MASK;
        acc >>= LIMB_BITS;
    }
    return r;
}

// a + b, limb-wise: a "wide" value (limbs < 2^30) usable as a mul operand.
template <class C>
HD Fp<C> add_w(const Fp<C>& a, const Fp<C>& b) {
    Fp<C> r;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) r.l[i] = a.l[i] + b.l[i];
    return r;
}

// a + 4m - b, limb-wise (a, b normalized, b < 4m): wide, limbs < 2^31.
template <class C>
HD Fp<C> sub_w(const Fp<C>& a, const Fp<C>& b) {
    Fp<C> r;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) r.l[i] = a.l[i] + C::SUB4[i] - b.l[i];
    return r;
}

// a + 8m - b - c (a, b, c normalized, b + c < 8m), normalized, < a + 8m.
template <class C>
HD Fp<C> sub2_n(const Fp<C>& a, const Fp<C>& b, const Fp<C>& c) {
    Fp<C> r;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) r.l[i] = a.l[i] + C::SUB8[i] - b.l[i] - c.l[i];
    return norm(r);
}

// a + 8m - b - c - d (all normalized, b + c + d < 8m), normalized, < a + 8m:
// the big borrow form of 8m keeps limbs 0..7 at >= 2^31 - 4 > 3 (2^29 - 1).
template <class C>
HD Fp<C> sub3_n(const Fp<C>& a, const Fp<C>& b, const Fp<C>& c, const Fp<C>& d) {
    Fp<C> r;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) r.l[i] = a.l[i] + C::SUB8[i] - b.l[i] - c.l[i] - d.l[i];
    return norm(r);
}

template <class C>
HD Fp<C> add(const Fp<C>& a, const Fp<C>& b) { return reduce(norm(add_w(a, b))); }

template <class C>
HD Fp<C> sub(const Fp<C>& a, const Fp<C>& b) { return reduce(norm(sub_w(a, b))); }

template <class C>
HD Fp<C> neg(const Fp<C>& a) { return reduce(norm(sub_w(fp_zero<C>(), a))); }

template <class C>
HD Fp<C> dbl(const Fp<C>& a) { return add(a, a); }

// Fully reduce a normalized value < 2m into [0, m).
template <class C>
HD Fp<C> canon(const Fp<C>& x) {
    Fp<C> d;
    int32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) {
        const int32_t v = int32_t(x.l[i]) - int32_t(C::MOD[i]) + borrow;
        d.l[i] = uint32_t(v) & LIMB_MASK;
        borrow = v >> LIMB_BITS;  // arithmetic shift: 0 or -1
    }
    const bool keep = borrow < 0;  // x < m
    Fp<C> r;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) r.l[i] = keep ? x.l[i] : d.l[i];
    return r;
}

// x == 0 mod m for normalized x < 2m  (x is 0 or m).
template <class C>
HD bool is_zero(const Fp<C>& x) {
    uint32_t z = 0, e = 0;
#pragma unroll
    for (int i = 0; i < LIMBS; ++i) {
        z |= x.l[i];
        e |= x.l[i] ^ C::MOD[i];
    }
    return z == 0 || e == 0;
}

template <class C>
HD bool eq(const Fp<C>& a, const Fp<C>& b) { return is_zero(sub(a, b)); }

// ---- memory form ---------------------------------------------------------

// Eight 32-bit words (value < 2^256) -> nine 29-bit limbs.
template <class C>
HD Fp<C> unpack(const uint32_t (&w)[8]) {
    Fp<C> r;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) {
        const int bit = LIMB_BITS * k, j = bit >> 5, s = bit & 31;
        uint32_t v = w[j] >> s;
        if (s > 32 - LIMB_BITS && j + 1 < 8) v |= w[j + 1] << (32 - s);
        r.l[k] = v & LIMB_MASK;
    }
    return r;
}

// Nine normalized limbs with value < 2^256 -> eight 32-bit words.
template <class C>
HD void pack(const Fp<C>& x, uint32_t (&w)[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int bit = 32 * j, k = bit / LIMB_BITS, s = bit % LIMB_BITS;
        uint32_t v = x.l[k] >> s;
        if (k + 1 < LIMBS) v |= x.l[k + 1] << (LIMB_BITS - s);
        if (k + 2 < LIMBS && s + 32 > 2 * LIMB_BITS) v |= x.l[k + 2] << (2 * LIMB_BITS - s);
        w[j] = v;
    }
}

// Canonical integer (< m, checked by the caller) -> Montgomery form (< 2m).
template <class C>
HD Fp<C> to_mont(const uint32_t (&w)[8]) {
    return mul(unpack<C>(w), fp_const<C>(C::R2));
}

// Montgomery form -> canonical integer in [0, m), eight words.
template <class C>
HD void from_mont(const Fp<C>& x, uint32_t (&w)[8]) {
    Fp<C> one = fp_zero<C>();
    one.l[0] = 1;
    pack(canon(mul(x, one)), w);
}

// Canonical-encoding check of eight LE words: value < m.
template <class C>
HD bool is_canonical(const uint32_t (&w)[8]) {
    bool lt = false, decided = false;
#pragma unroll
    for (int i = 7; i >= 0; --i) {
        if (!decided && w[i] != C::MOD32[i]) {
            lt = w[i] < C::MOD32[i];
            decided = true;
        }
    }
    return decided && lt;
}

// a^e for a small public exponent given as 64-bit words, high to low.
template <class C, int W>
HD Fp<C> pow_words(const Fp<C>& a, const uint64_t (&e)[W]) {
    Fp<C> r = fp_one<C>();
    bool started = false;
    for (int w = W - 1; w >= 0; --w) {
        for (int b = 63; b >= 0; --b) {
            if (started) r = sqr(r);
            if ((e[w] >> b) & 1) {
                r = started ? mul(r, a) : a;
                started = true;
            }
        }
    }
    return started ? r : fp_one<C>();
}

// Inverse by Fermat: a^(m-2); a must be nonzero mod m.
template <class C>
HD Fp<C> inv(const Fp<C>& a) {
    uint64_t e[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        e[i] = uint64_t(C::MOD32[2 * i]) | (uint64_t(C::MOD32[2 * i + 1]) << 32);
    e[0] -= 2;  // m is odd and > 2: no borrow
    return pow_words<C, 4>(a, e);
}

// ---- Fq2 = Fq[u]/(u^2 + 1) -------------------------------------------------

struct Fq2 {
    Fq c0, c1;
};

HD Fq2 f2_zero() { return {fp_zero<FqParams>(), fp_zero<FqParams>()}; }
HD Fq2 f2_one() { return {fp_one<FqParams>(), fp_zero<FqParams>()}; }
HD Fq2 f2_add(const Fq2& a, const Fq2& b) { return {add(a.c0, b.c0), add(a.c1, b.c1)}; }
HD Fq2 f2_sub(const Fq2& a, const Fq2& b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1)}; }
HD Fq2 f2_neg(const Fq2& a) { return {neg(a.c0), neg(a.c1)}; }
HD Fq2 f2_dbl(const Fq2& a) { return f2_add(a, a); }
HD Fq2 f2_conj(const Fq2& a) { return {a.c0, neg(a.c1)}; }

// Karatsuba: three Fq products.
HD Fq2 f2_mul(const Fq2& a, const Fq2& b) {
    const Fq t0 = mul(a.c0, b.c0);
    const Fq t1 = mul(a.c1, b.c1);
    const Fq t2 = mul(add_w(a.c0, a.c1), norm(add_w(b.c0, b.c1)));
    // c0 = t0 - t1 ; c1 = t2 - t0 - t1
    return {sub(t0, t1), reduce(sub2_n(t2, t0, t1))};
}

HD Fq2 f2_sqr(const Fq2& a) {
    // (a0 + a1)(a0 - a1), 2 a0 a1
    const Fq s = mul(add_w(a.c0, a.c1), norm(sub_w(a.c0, a.c1)));
    const Fq t = mul(a.c0, a.c1);
    return {s, add(t, t)};
}

HD Fq2 f2_mul_fq(const Fq2& a, const Fq& b) { return {mul(a.c0, b), mul(a.c1, b)}; }

HD bool f2_is_zero(const Fq2& a) { return is_zero(a.c0) && is_zero(a.c1); }
HD bool f2_eq(const Fq2& a, const Fq2& b) { return eq(a.c0, b.c0) && eq(a.c1, b.c1); }

HD Fq2 f2_inv(const Fq2& a) {
    const Fq n = add(sqr(a.c0), sqr(a.c1));
    const Fq ni = inv(n);
    return {mul(a.c0, ni), neg(mul(a.c1, ni))};
}

HD Fq2 f2_canon(const Fq2& a) { return {canon(a.c0), canon(a.c1)}; }

}  // namespace bn254
