// BN254 G1 arithmetic for the host epilogues (the window combination of an
// MSM: a few hundred dependent doublings and additions on one core).
//
// Field elements are four 64-bit limbs in Montgomery form with R = 2^256,
// fully reduced; products are CIOS on unsigned __int128 (about four times
// fewer multiply instructions than the device's 29-bit limbs compiled for the
// host).  Device words (eight 32-bit words, Montgomery with R = 2^261, below
// 2^256) enter by one product with 2^251: x 2^261 2^251 2^-256 = x 2^256.
// Points are XYZZ (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2), identity ZZ = 0, with
// the same formulas as the device (dbl-2008-s-1, add-2008-s, a = 0).
#pragma once
#include <stdint.h>

#include "bn254_constants.hpp"

namespace dvgpu {
namespace hostg1 {

using u128 = unsigned __int128;

struct Fe {
    uint64_t l[4];
};

inline constexpr uint64_t word64(const uint32_t* w, int i) { return uint64_t(w[2 * i]) | (uint64_t(w[2 * i + 1]) << 32); }

constexpr uint64_t P[4] = {word64(bn254::FqParams::MOD32, 0), word64(bn254::FqParams::MOD32, 1),
                           word64(bn254::FqParams::MOD32, 2), word64(bn254::FqParams::MOD32, 3)};

// -p^-1 mod 2^64 by Newton iteration (each step doubles the correct bits).
constexpr uint64_t neg_inv64(uint64_t p0) {
    uint64_t x = 1;
    for (int i = 0; i < 6; ++i) x *= 2 - p0 * x;
    return ~x + 1;
}
constexpr uint64_t PINV = neg_inv64(P[0]);

// r - p when r >= p, else r (r < 2p), without branches.
inline void reduce_once(uint64_t (&r)[4]) {
    uint64_t s[4], borrow = 0;
    for (int i = 0; i < 4; ++i) {
        const u128 d = u128(r[i]) - P[i] - borrow;
        s[i] = uint64_t(d);
        borrow = uint64_t(d >> 64) & 1;
    }
    const uint64_t keep = 0 - borrow;  // all ones when r < p
    for (int i = 0; i < 4; ++i) r[i] = (r[i] & keep) | (s[i] & ~keep);
}

// Montgomery product by CIOS.  The top limb of p is below 2^62, so the
// running value stays below 2p < 2^255 and the textbook form's extra carry
// words are never needed (the high word of each step's two carries is
// folded into the top limb directly).
inline Fe mul(const Fe& a, const Fe& b) {
    uint64_t t[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        const uint64_t bi = b.l[i];
        u128 A = u128(a.l[0]) * bi + t[0];
        const uint64_t lo = uint64_t(A);
        const uint64_t m = lo * PINV;
        u128 C = (u128(m) * P[0] + lo) >> 64;  // the low word cancels
        A >>= 64;
        for (int j = 1; j < 4; ++j) {
            A += u128(a.l[j]) * bi + t[j];
            const uint64_t tj = uint64_t(A);
            A >>= 64;
            C += u128(m) * P[j] + tj;
            t[j - 1] = uint64_t(C);
            C >>= 64;
        }
        t[3] = uint64_t(C) + uint64_t(A);
    }
    uint64_t r[4] = {t[0], t[1], t[2], t[3]};
    reduce_once(r);
    return {{r[0], r[1], r[2], r[3]}};
}

// a + b for a, b < p: the sum is below 2^255, so one subtraction suffices.
inline Fe add(const Fe& a, const Fe& b) {
    uint64_t r[4];
    uint64_t c = 0;
    for (int i = 0; i < 4; ++i) {
        const u128 s = u128(a.l[i]) + b.l[i] + c;
        r[i] = uint64_t(s);
        c = uint64_t(s >> 64);
    }
    reduce_once(r);
    return {{r[0], r[1], r[2], r[3]}};
}

inline Fe sub(const Fe& a, const Fe& b) {
    uint64_t r[4], borrow = 0;
    for (int i = 0; i < 4; ++i) {
        const u128 d = u128(a.l[i]) - b.l[i] - borrow;
        r[i] = uint64_t(d);
        borrow = uint64_t(d >> 64) & 1;
    }
    const uint64_t mask = 0 - borrow;  // add p back when the difference went negative
    uint64_t c = 0;
    for (int i = 0; i < 4; ++i) {
        const u128 s = u128(r[i]) + (P[i] & mask) + c;
        r[i] = uint64_t(s);
        c = uint64_t(s >> 64);
    }
    return {{r[0], r[1], r[2], r[3]}};
}

inline bool is_zero(const Fe& a) { return !(a.l[0] | a.l[1] | a.l[2] | a.l[3]); }

// Device Montgomery words (R = 2^261, any value below 2^256) -> host form.
inline Fe from_device(const uint32_t* w) {
    const Fe x = {{word64(w, 0), word64(w, 1), word64(w, 2), word64(w, 3)}};
    const Fe c = {{0, 0, 0, uint64_t(1) << (251 - 192)}};  // 2^251 < p
    return mul(x, c);
}

// Host form -> canonical little-endian 32-bit words.
inline void to_words(const Fe& a, uint32_t (&w)[8]) {
    const Fe x = mul(a, Fe{{1, 0, 0, 0}});
    for (int i = 0; i < 4; ++i) {
        w[2 * i] = uint32_t(x.l[i]);
        w[2 * i + 1] = uint32_t(x.l[i] >> 32);
    }
}

// a^(p - 2) by left-to-right square and multiply (a nonzero).
inline Fe inv(const Fe& a) {
    uint64_t e[4] = {P[0] - 2, P[1], P[2], P[3]};  // p is odd and > 2: no borrow
    Fe r = a;
    bool started = false;
    for (int i = 3; i >= 0; --i)
        for (int b = 63; b >= 0; --b) {
            if (started) r = mul(r, r);
            if ((e[i] >> b) & 1) {
                r = started ? mul(r, a) : a;
                started = true;
            }
        }
    return r;
}

struct Xyzz {
    Fe X, Y, ZZ, ZZZ;
};

inline Xyzz infinity() { return {{{0, 0, 0, 0}}, {{0, 0, 0, 0}}, {{0, 0, 0, 0}}, {{0, 0, 0, 0}}}; }
inline bool is_inf(const Xyzz& p) { return is_zero(p.ZZ); }

// Device XYZZ words (X || Y || ZZ || ZZZ, eight words each) -> host point.
inline Xyzz from_device_xyzz(const uint32_t* w) {
    Xyzz p = {from_device(w), from_device(w + 8), from_device(w + 16), from_device(w + 24)};
    if (is_zero(p.ZZ)) p = infinity();
    return p;
}

inline Xyzz dbl(const Xyzz& p) {
    if (is_inf(p) || is_zero(p.Y)) return infinity();
    const Fe U = add(p.Y, p.Y);
    const Fe V = mul(U, U);
    const Fe W = mul(U, V);
    const Fe S = mul(p.X, V);
    const Fe xx = mul(p.X, p.X);
    const Fe M = add(add(xx, xx), xx);
    const Fe X3 = sub(sub(mul(M, M), S), S);
    const Fe Y3 = sub(mul(M, sub(S, X3)), mul(W, p.Y));
    return {X3, Y3, mul(V, p.ZZ), mul(W, p.ZZZ)};
}

inline Xyzz add(const Xyzz& p, const Xyzz& q) {
    if (is_inf(p)) return q;
    if (is_inf(q)) return p;
    const Fe U1 = mul(p.X, q.ZZ), U2 = mul(q.X, p.ZZ);
    const Fe S1 = mul(p.Y, q.ZZZ), S2 = mul(q.Y, p.ZZZ);
    const Fe Pd = sub(U2, U1), R = sub(S2, S1);
    if (is_zero(Pd)) return is_zero(R) ? dbl(p) : infinity();
    const Fe PP = mul(Pd, Pd), PPP = mul(Pd, PP), Q = mul(U1, PP);
    const Fe X3 = sub(sub(sub(mul(R, R), PPP), Q), Q);
    const Fe Y3 = sub(mul(R, sub(Q, X3)), mul(S1, PPP));
    return {X3, Y3, mul(mul(p.ZZ, q.ZZ), PP), mul(mul(p.ZZZ, q.ZZZ), PPP)};
}

// n >= 1 doublings.  They run in Jacobian coordinates (dbl-2009-l for
// a = 0: 2M + 5S and a shorter dependency chain than the XYZZ doubling),
// entered and left in XYZZ: (X, Y, ZZ, ZZZ) is the Jacobian point
// (X ZZ ZZZ^2, Y ZZ^3 ZZZ^2, ZZ ZZZ), and Jacobian (X, Y, Z) is the XYZZ
// point (X, Y, Z^2, Z^3).  The identity (Z = 0) stays the identity.
inline Xyzz dbl_n(const Xyzz& p, int n) {
    if (is_inf(p)) return p;
    const Fe t = mul(p.ZZZ, p.ZZZ), zz2 = mul(p.ZZ, p.ZZ);
    Fe X = mul(mul(p.X, p.ZZ), t);
    Fe Y = mul(mul(p.Y, mul(zz2, p.ZZ)), t);
    Fe Z = mul(p.ZZ, p.ZZZ);
    for (int k = 0; k < n; ++k) {
        const Fe A = mul(X, X), B = mul(Y, Y), C = mul(B, B);
        const Fe xb = add(X, B);
        const Fe D2 = sub(sub(mul(xb, xb), A), C);
        const Fe D = add(D2, D2);
        const Fe E = add(add(A, A), A);
        const Fe F = mul(E, E);
        const Fe X3 = sub(F, add(D, D));
        const Fe C2 = add(C, C), C4 = add(C2, C2), C8 = add(C4, C4);
        const Fe Y3 = sub(mul(E, sub(D, X3)), C8);
        const Fe yz = mul(Y, Z);
        Z = add(yz, yz);
        X = X3;
        Y = Y3;
    }
    if (is_zero(Z)) return infinity();
    const Fe ZZ = mul(Z, Z);
    return {X, Y, ZZ, mul(ZZ, Z)};
}

// Canonical affine words (x || y, zero words for the identity).
inline void to_affine_words(const Xyzz& p, uint32_t (&x)[8], uint32_t (&y)[8]) {
    if (is_inf(p)) {
        for (int i = 0; i < 8; ++i) x[i] = y[i] = 0;
        return;
    }
    const Fe t = inv(mul(p.ZZ, p.ZZZ));
    to_words(mul(p.X, mul(t, p.ZZZ)), x);  // X / ZZ
    to_words(mul(p.Y, mul(t, p.ZZ)), y);   // Y / ZZZ
}

}  // namespace hostg1
}  // namespace dvgpu
