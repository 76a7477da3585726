// GLV scalar multiplication on BN254 G1 (Gallant, Lambert, Vanstone,
// "Faster point multiplication on elliptic curves with efficient
// endomorphisms", CRYPTO 2001).
//
// phi(x, y) = (beta x, y) acts on G1 as [lambda] with lambda^2 + lambda + 1
// = 0 mod r.  A scalar k < r splits as k = k1 + k2 lambda (mod r) with
// |k1|, |k2| < 2^126 by rounding against the short lattice basis
// (a1, -a2), (a2, b2) of {(a, b) : a + b lambda = 0 mod r}:
//   c1 = round(b2 k / r), c2 = round(a2 k / r),
//   k1 = k - c1 a1 - c2 a2,  k2 = c1 a2 - c2 b2.
// The two roundings use the truncated reciprocals floor(2^384 b2 / r) and
// floor(2^384 a2 / r); their error is below 2^-130, so each |c - exact| is
// at most 1/2 + 2^-130 and the bound 2^126 holds for every k < r.
// [k]P is then one joint double-and-add over 126 bit positions on the table
// {P1, P2, P1 + P2} with P1 = [sign k1]P, P2 = [sign k2]phi(P), or, for a
// scalar shared by the wave, over the joint sparse form of (|k1|, |k2|).
#pragma once
#include "curve.hpp"

namespace bn254 {

struct GlvSplit {
    uint32_t k1[4], k2[4];  // magnitudes, little-endian words (< 2^126)
    uint32_t neg1, neg2;    // 1 when k1 (k2) is negative
};

constexpr int GLV_BITS = 126;

// out[0 .. NA+NB) = a * b
template <int NA, int NB>
HD void mp_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) {
    for (int i = 0; i < NA + NB; ++i) out[i] = 0;
    for (int i = 0; i < NA; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < NB; ++j) {
            const uint64_t t = uint64_t(a[i]) * b[j] + out[i + j] + carry;
            out[i + j] = uint32_t(t);
            carry = t >> 32;
        }
        out[i + NB] = uint32_t(carry);
    }
}

// acc (N words, two's complement) -= v (M <= N words)
template <int N, int M>
HD void mp_sub(uint32_t* acc, const uint32_t* v) {
    uint64_t borrow = 0;
    for (int i = 0; i < N; ++i) {
        const uint64_t d = uint64_t(acc[i]) - (i < M ? v[i] : 0u) - borrow;
        acc[i] = uint32_t(d);
        borrow = (d >> 32) & 1;
    }
}

// |v| of an N-word two's complement value into 4 words; returns the sign.
template <int N>
HD uint32_t mp_abs4(const uint32_t* v, uint32_t (&out)[4]) {
    const uint32_t neg = v[N - 1] >> 31;
    uint64_t carry = neg;
    for (int i = 0; i < 4; ++i) {
        const uint64_t t = uint64_t(neg ? ~v[i] : v[i]) + carry;
        out[i] = uint32_t(t);
        carry = t >> 32;
    }
    return neg;
}

// round(k * rc / 2^384) for an NC-word reciprocal; the quotient has W words.
template <int NC, int W>
HD void glv_round(const uint32_t (&k)[8], const uint32_t (&rc)[NC], uint32_t (&c)[W]) {
    uint32_t p[8 + NC];
    mp_mul<8, NC>(k, rc, p);
    uint64_t t = uint64_t(p[11]) + 0x80000000u;  // + 2^383
    p[11] = uint32_t(t);
    for (int i = 12; i < 8 + NC; ++i) {
        t = uint64_t(p[i]) + (t >> 32);
        p[i] = uint32_t(t);
    }
    for (int i = 0; i < W; ++i) c[i] = 12 + i < 8 + NC ? p[12 + i] : 0u;
}

// k: canonical little-endian words, k < r.
HD GlvSplit glv_split(const uint32_t (&k)[8]) {
    uint32_t c1[4], c2[3];
    glv_round<9, 4>(k, GLV_RC1, c1);  // round(b2 k / r) < 2^128
    glv_round<7, 3>(k, GLV_RC2, c2);  // round(a2 k / r) < 2^65
    uint32_t t[9] = {k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7], 0};
    uint32_t u[8];
    mp_mul<4, 4>(c1, GLV_A1, u);
    mp_sub<9, 8>(t, u);
    uint32_t v[5];
    mp_mul<3, 2>(c2, GLV_A2, v);
    mp_sub<9, 5>(t, v);
    GlvSplit s;
    s.neg1 = mp_abs4<9>(t, s.k1);
    uint32_t w[9];
    mp_mul<4, 2>(c1, GLV_A2, w);  // c1 a2 (6 words)
    w[6] = w[7] = w[8] = 0;
    uint32_t x[7];
    mp_mul<3, 4>(c2, GLV_B2, x);  // c2 b2
    mp_sub<9, 7>(w, x);
    s.neg2 = mp_abs4<9>(w, s.k2);
    return s;
}

// Joint sparse form of (a, b), both < 2^126 (Solinas, "Low-weight binary
// representations for pairs of integers", 2001; Hankerson-Menezes-Vanstone,
// Guide to ECC, Algorithm 3.50): digits u1, u2 in {-1, 0, 1} with
// a = sum u1_l 2^l, b = sum u2_l 2^l, at most 127 positions, and on average
// half of the columns (u1_l, u2_l) nonzero.  Bit l of nz/sg marks digit l
// nonzero / negative.  a and b are consumed.
struct Jsf {
    uint32_t nz1[4], sg1[4], nz2[4], sg2[4];
};

HD Jsf jsf_recode(uint32_t (&a)[4], uint32_t (&b)[4]) {
    Jsf j = {};
    uint32_t d1 = 0, d2 = 0;
    for (int l = 0; l < 4 * 32; ++l) {
        const uint32_t l1 = (a[0] + d1) & 7, l2 = (b[0] + d2) & 7;
        int u1 = 0, u2 = 0;
        if (l1 & 1) {
            u1 = (l1 & 3) == 1 ? 1 : -1;
            if ((l1 == 3 || l1 == 5) && (l2 & 3) == 2) u1 = -u1;
        }
        if (l2 & 1) {
            u2 = (l2 & 3) == 1 ? 1 : -1;
            if ((l2 == 3 || l2 == 5) && (l1 & 3) == 2) u2 = -u2;
        }
        const uint32_t bit = 1u << (l & 31);
        if (u1) j.nz1[l >> 5] |= bit;
        if (u1 < 0) j.sg1[l >> 5] |= bit;
        if (u2) j.nz2[l >> 5] |= bit;
        if (u2 < 0) j.sg2[l >> 5] |= bit;
        if (2 * int(d1) == 1 + u1) d1 = 1 - d1;
        if (2 * int(d2) == 1 + u2) d2 = 1 - d2;
        for (int w = 0; w < 3; ++w) {
            a[w] = (a[w] >> 1) | (a[w + 1] << 31);
            b[w] = (b[w] >> 1) | (b[w + 1] << 31);
        }
        a[3] >>= 1;
        b[3] >>= 1;
    }
    return j;
}

}  // namespace bn254

namespace dvgpu {
using bn254::GlvSplit;

__device__ __forceinline__ uint32_t glv_digit(const uint32_t (&k1)[4], const uint32_t (&k2)[4], int b) {
    return ((k1[b >> 5] >> (b & 31)) & 1u) | (((k2[b >> 5] >> (b & 31)) & 1u) << 1);
}

// [k]P for a per-lane split.  Every step doubles and adds the selected table
// entry; lanes whose digit is zero keep the doubled value.
__device__ inline bn254::Xyzz<bn254::FqOps> g1_mul_glv(const bn254::Xyzz<bn254::FqOps>& p, const GlvSplit& s) {
    using namespace bn254;
    using P = Xyzz<FqOps>;
    const Fq bx = mul(p.X, fp_const<FqParams>(GLV_BETA));
    const Fq ny = neg(p.Y);
    const P p1 = {p.X, s.neg1 ? ny : p.Y, p.ZZ, p.ZZZ};
    const P p2 = {bx, s.neg2 ? ny : p.Y, p.ZZ, p.ZZZ};
    const P sum = xyzz_add<FqOps>(p1, p2);
    P acc = xyzz_inf<FqOps>();
    for (int b = GLV_BITS - 1; b >= 0; --b) {
        acc = xyzz_dbl<FqOps>(acc);
        const uint32_t d = glv_digit(s.k1, s.k2, b);
        if (d) acc = xyzz_add<FqOps>(acc, d == 3 ? sum : (d == 2 ? p2 : p1));
    }
    return acc;
}

// Same result for a split that is identical in every active lane: the split
// is recoded to joint sparse form in scalar registers and the table grows by
// P1 - P2, so about half of the 127 steps add (three quarters with plain
// joint bits) and every branch is on a scalar digit.
__device__ inline bn254::Xyzz<bn254::FqOps> g1_mul_glv_uniform(const bn254::Xyzz<bn254::FqOps>& p,
                                                                const GlvSplit& s) {
    using namespace bn254;
    using P = Xyzz<FqOps>;
    uint32_t k1[4], k2[4];
    for (int i = 0; i < 4; ++i) {
        k1[i] = __builtin_amdgcn_readfirstlane(s.k1[i]);
        k2[i] = __builtin_amdgcn_readfirstlane(s.k2[i]);
    }
    const uint32_t n1 = __builtin_amdgcn_readfirstlane(s.neg1), n2 = __builtin_amdgcn_readfirstlane(s.neg2);
    const Jsf j = jsf_recode(k1, k2);
    const Fq bx = mul(p.X, fp_const<FqParams>(GLV_BETA));
    const Fq ny = neg(p.Y);
    const Fq y1 = n1 ? ny : p.Y, y2 = n2 ? ny : p.Y;
    const P sum = xyzz_add<FqOps>({p.X, y1, p.ZZ, p.ZZZ}, {bx, y2, p.ZZ, p.ZZZ});
    const P dif = xyzz_add<FqOps>({p.X, y1, p.ZZ, p.ZZZ}, {bx, neg(y2), p.ZZ, p.ZZZ});
    P acc = xyzz_inf<FqOps>();
    for (int l = 4 * 32 - 1; l >= 0; --l) {
        acc = xyzz_dbl<FqOps>(acc);
        const uint32_t w = l >> 5, bit = 1u << (l & 31);
        const bool a1 = j.nz1[w] & bit, a2 = j.nz2[w] & bit;
        if (!(a1 || a2)) continue;
        const bool s1 = j.sg1[w] & bit, s2 = j.sg2[w] & bit;
        P t;
        bool negate;
        if (a1 && a2) {
            t = s1 == s2 ? sum : dif;  // u1 P1 + u2 P2 = +-(P1 + P2) or +-(P1 - P2)
            negate = s1;
        } else if (a1) {
            t = {p.X, y1, p.ZZ, p.ZZZ};
            negate = s1;
        } else {
            t = {bx, y2, p.ZZ, p.ZZZ};
            negate = s2;
        }
        if (negate) t.Y = neg(t.Y);
        acc = xyzz_add<FqOps>(acc, t);
    }
    return acc;
}

// [k]P for canonical scalar words; `uniform` when the scalar is the same in
// every active lane of the wave.
__device__ inline bn254::Xyzz<bn254::FqOps> g1_mul_scalar(const bn254::Xyzz<bn254::FqOps>& p,
                                                           const uint32_t (&k)[8], bool uniform) {
    const GlvSplit s = bn254::glv_split(k);
    return uniform ? g1_mul_glv_uniform(p, s) : g1_mul_glv(p, s);
}
}  // namespace dvgpu
