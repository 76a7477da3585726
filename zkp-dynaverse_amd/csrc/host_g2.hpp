// BN254 G2 arithmetic for the host epilogues (the window combination of a
// G2 MSM: a few hundred dependent doublings and additions on one core), over
// the 64-bit-limb Fq of host_g1.hpp.
//
// Fq2 = Fq[u]/(u^2 + 1), products by Karatsuba (three Fq products).  Points
// are XYZZ over Fq2 with the formulas of host_g1.hpp (the twist has a = 0
// too; b' never enters an addition or doubling).  Device words: an Fq2 is
// c0 || c1, eight Montgomery words each (R = 2^261), a point X || Y || ZZ ||
// ZZZ, 64 words.
#pragma once
#include "host_g1.hpp"

namespace dvgpu {
namespace hostg2 {

using hostg1::Fe;

struct F2 {
    Fe c0, c1;
};

inline F2 add(const F2& a, const F2& b) { return {hostg1::add(a.c0, b.c0), hostg1::add(a.c1, b.c1)}; }
inline F2 sub(const F2& a, const F2& b) { return {hostg1::sub(a.c0, b.c0), hostg1::sub(a.c1, b.c1)}; }
inline F2 mul(const F2& a, const F2& b) {
    const Fe t0 = hostg1::mul(a.c0, b.c0), t1 = hostg1::mul(a.c1, b.c1);
    const Fe t2 = hostg1::mul(hostg1::add(a.c0, a.c1), hostg1::add(b.c0, b.c1));
    return {hostg1::sub(t0, t1), hostg1::sub(hostg1::sub(t2, t0), t1)};
}
inline bool is_zero(const F2& a) { return hostg1::is_zero(a.c0) && hostg1::is_zero(a.c1); }

// a^-1 = conj(a) / (c0^2 + c1^2) (a nonzero).
inline F2 inv(const F2& a) {
    const Fe n = hostg1::inv(hostg1::add(hostg1::mul(a.c0, a.c0), hostg1::mul(a.c1, a.c1)));
    return {hostg1::mul(a.c0, n), hostg1::sub(Fe{{0, 0, 0, 0}}, hostg1::mul(a.c1, n))};
}

struct Xyzz {
    F2 X, Y, ZZ, ZZZ;
};

inline Xyzz infinity() {
    const F2 z = {{{0, 0, 0, 0}}, {{0, 0, 0, 0}}};
    return {z, z, z, z};
}
inline bool is_inf(const Xyzz& p) { return is_zero(p.ZZ); }

inline F2 from_device(const uint32_t* w) { return {hostg1::from_device(w), hostg1::from_device(w + 8)}; }

// Device XYZZ words (64) -> host point.
inline Xyzz from_device_xyzz(const uint32_t* w) {
    Xyzz p = {from_device(w), from_device(w + 16), from_device(w + 32), from_device(w + 48)};
    if (is_zero(p.ZZ)) p = infinity();
    return p;
}

inline Xyzz dbl(const Xyzz& p) {
    if (is_inf(p) || is_zero(p.Y)) return infinity();
    const F2 U = add(p.Y, p.Y);
    const F2 V = mul(U, U);
    const F2 W = mul(U, V);
    const F2 S = mul(p.X, V);
    const F2 xx = mul(p.X, p.X);
    const F2 M = add(add(xx, xx), xx);
    const F2 X3 = sub(sub(mul(M, M), S), S);
    const F2 Y3 = sub(mul(M, sub(S, X3)), mul(W, p.Y));
    return {X3, Y3, mul(V, p.ZZ), mul(W, p.ZZZ)};
}

inline Xyzz add(const Xyzz& p, const Xyzz& q) {
    if (is_inf(p)) return q;
    if (is_inf(q)) return p;
    const F2 U1 = mul(p.X, q.ZZ), U2 = mul(q.X, p.ZZ);
    const F2 S1 = mul(p.Y, q.ZZZ), S2 = mul(q.Y, p.ZZZ);
    const F2 Pd = sub(U2, U1), R = sub(S2, S1);
    if (is_zero(Pd)) return is_zero(R) ? dbl(p) : infinity();
    const F2 PP = mul(Pd, Pd), PPP = mul(Pd, PP), Q = mul(U1, PP);
    const F2 X3 = sub(sub(sub(mul(R, R), PPP), Q), Q);
    const F2 Y3 = sub(mul(R, sub(Q, X3)), mul(S1, PPP));
    return {X3, Y3, mul(mul(p.ZZ, q.ZZ), PP), mul(mul(p.ZZZ, q.ZZZ), PPP)};
}

// n >= 1 doublings in Jacobian coordinates (dbl-2009-l, a = 0), entered
// and left in XYZZ as in host_g1.hpp.
inline Xyzz dbl_n(const Xyzz& p, int n) {
    if (is_inf(p)) return p;
    const F2 t = mul(p.ZZZ, p.ZZZ), zz2 = mul(p.ZZ, p.ZZ);
    F2 X = mul(mul(p.X, p.ZZ), t);
    F2 Y = mul(mul(p.Y, mul(zz2, p.ZZ)), t);
    F2 Z = mul(p.ZZ, p.ZZZ);
    for (int k = 0; k < n; ++k) {
        const F2 A = mul(X, X), B = mul(Y, Y), C = mul(B, B);
        const F2 xb = add(X, B);
        const F2 D2 = sub(sub(mul(xb, xb), A), C);
        const F2 D = add(D2, D2);
        const F2 E = add(add(A, A), A);
        const F2 F = mul(E, E);
        const F2 X3 = sub(F, add(D, D));
        const F2 C2 = add(C, C), C4 = add(C2, C2), C8 = add(C4, C4);
        const F2 Y3 = sub(mul(E, sub(D, X3)), C8);
        const F2 yz = mul(Y, Z);
        Z = add(yz, yz);
        X = X3;
        Y = Y3;
    }
    if (is_zero(Z)) return infinity();
    const F2 ZZ = mul(Z, Z);
    return {X, Y, ZZ, mul(ZZ, Z)};
}

// Canonical affine words x.c0, x.c1, y.c0, y.c1 (zero words for the identity).
inline void to_affine_words(const Xyzz& p, uint32_t (&w)[4][8]) {
    if (is_inf(p)) {
        for (int k = 0; k < 4; ++k)
            for (int i = 0; i < 8; ++i) w[k][i] = 0;
        return;
    }
    const F2 t = inv(mul(p.ZZ, p.ZZZ));
    const F2 x = mul(p.X, mul(t, p.ZZZ));  // X / ZZ
    const F2 y = mul(p.Y, mul(t, p.ZZ));   // Y / ZZZ
    hostg1::to_words(x.c0, w[0]);
    hostg1::to_words(x.c1, w[1]);
    hostg1::to_words(y.c0, w[2]);
    hostg1::to_words(y.c1, w[3]);
}

}  // namespace hostg2
}  // namespace dvgpu
