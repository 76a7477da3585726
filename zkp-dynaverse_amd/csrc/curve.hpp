// BN254 G1 (over Fq) and G2 (over Fq2) in XYZZ coordinates:
// x = X/ZZ, y = Y/ZZZ with ZZ^3 = ZZZ^2; infinity has ZZ = ZZZ = 0.
// Formulas: EFD "xyzz" short-Weierstrass a = 0 family
//   madd-2008-s (8M+2S), add-2008-s (12M+2S), dbl-2008-s-1, mdbl-2008-s-1,
// with explicit handling of the exceptional cases (P == Q, P == -Q, either
// operand at infinity) so that every result is exact.
#pragma once
#include "field.hpp"

namespace bn254 {

// Field-operation traits so one set of formulas serves G1 and G2.
struct FqOps {
    using E = Fq;
    static HD E zero() { return fp_zero<FqParams>(); }
    static HD E one() { return fp_one<FqParams>(); }
    static HD E mul(const E& a, const E& b) { return bn254::mul(a, b); }
    static HD E sqr(const E& a) { return bn254::sqr(a); }
    static HD E add(const E& a, const E& b) { return bn254::add(a, b); }
    static HD E sub(const E& a, const E& b) { return bn254::sub(a, b); }
    static HD E neg(const E& a) { return bn254::neg(a); }
    static HD bool is_zero(const E& a) { return bn254::is_zero(a); }
    static HD E inv(const E& a) { return bn254::inv(a); }
    static HD E canon(const E& a) { return bn254::canon(a); }
    static HD E b() { return fp_const<FqParams>(FQ_B3); }
};

struct Fq2Ops {
    using E = Fq2;
    static HD E zero() { return f2_zero(); }
    static HD E one() { return f2_one(); }
    static HD E mul(const E& a, const E& b) { return f2_mul(a, b); }
    static HD E sqr(const E& a) { return f2_sqr(a); }
    static HD E add(const E& a, const E& b) { return f2_add(a, b); }
    static HD E sub(const E& a, const E& b) { return f2_sub(a, b); }
    static HD E neg(const E& a) { return f2_neg(a); }
    static HD bool is_zero(const E& a) { return f2_is_zero(a); }
    static HD E inv(const E& a) { return f2_inv(a); }
    static HD E canon(const E& a) { return f2_canon(a); }
    static HD E b() { return {fp_const<FqParams>(FQ2_B_C0), fp_const<FqParams>(FQ2_B_C1)}; }
};

template <class F>
struct Affine {
    typename F::E x, y;
    bool inf;
};

template <class F>
struct Xyzz {
    typename F::E X, Y, ZZ, ZZZ;
};

template <class F>
HD Xyzz<F> xyzz_inf() { return {F::zero(), F::zero(), F::zero(), F::zero()}; }

template <class F>
HD bool xyzz_is_inf(const Xyzz<F>& p) { return F::is_zero(p.ZZ); }

template <class F>
HD Xyzz<F> xyzz_from_affine(const Affine<F>& a) {
    if (a.inf) return xyzz_inf<F>();
    return {a.x, a.y, F::one(), F::one()};
}

template <class F>
HD Xyzz<F> xyzz_neg(const Xyzz<F>& p) { return {p.X, F::neg(p.Y), p.ZZ, p.ZZZ}; }

// 2P for an affine P = (x, y) (mdbl-2008-s-1).
template <class F>
HD Xyzz<F> xyzz_mdbl(const typename F::E& x, const typename F::E& y) {
    using E = typename F::E;
    if (F::is_zero(y)) return xyzz_inf<F>();
    const E U = F::add(y, y);
    const E V = F::sqr(U);
    const E W = F::mul(U, V);
    const E S = F::mul(x, V);
    const E xx = F::sqr(x);
    const E M = F::add(F::add(xx, xx), xx);
    const E X3 = F::sub(F::sqr(M), F::add(S, S));
    const E Y3 = F::sub(F::mul(M, F::sub(S, X3)), F::mul(W, y));
    return {X3, Y3, V, W};
}

// 2P (dbl-2008-s-1, a = 0).
template <class F>
HD Xyzz<F> xyzz_dbl(const Xyzz<F>& p) {
    using E = typename F::E;
    if (F::is_zero(p.ZZ) || F::is_zero(p.Y)) return xyzz_inf<F>();
    const E U = F::add(p.Y, p.Y);
    const E V = F::sqr(U);
    const E W = F::mul(U, V);
    const E S = F::mul(p.X, V);
    const E xx = F::sqr(p.X);
    const E M = F::add(F::add(xx, xx), xx);
    const E X3 = F::sub(F::sqr(M), F::add(S, S));
    const E Y3 = F::sub(F::mul(M, F::sub(S, X3)), F::mul(W, p.Y));
    return {X3, Y3, F::mul(V, p.ZZ), F::mul(W, p.ZZZ)};
}

// P + (x2, y2) with the affine operand finite (madd-2008-s).
template <class F>
HD Xyzz<F> xyzz_madd(const Xyzz<F>& p, const typename F::E& x2, const typename F::E& y2) {
    using E = typename F::E;
    if (F::is_zero(p.ZZ)) return {x2, y2, F::one(), F::one()};
    const E U2 = F::mul(x2, p.ZZ);
    const E S2 = F::mul(y2, p.ZZZ);
    const E P = F::sub(U2, p.X);
    const E R = F::sub(S2, p.Y);
    if (F::is_zero(P)) {
        if (F::is_zero(R)) return xyzz_mdbl<F>(x2, y2);
        return xyzz_inf<F>();
    }
    const E PP = F::sqr(P);
    const E PPP = F::mul(P, PP);
    const E Q = F::mul(p.X, PP);
    const E X3 = F::sub(F::sub(F::sqr(R), PPP), F::add(Q, Q));
    const E Y3 = F::sub(F::mul(R, F::sub(Q, X3)), F::mul(p.Y, PPP));
    return {X3, Y3, F::mul(p.ZZ, PP), F::mul(p.ZZZ, PPP)};
}

template <class F>
HD Xyzz<F> xyzz_madd(const Xyzz<F>& p, const Affine<F>& a) {
    if (a.inf) return p;
    return xyzz_madd<F>(p, a.x, a.y);
}

// P + (x2, y2) or P + (x2, -y2) (the signed-digit MSM's bucket accumulation);
// (x2, y2) finite and fully reduced.
template <class F>
HD Xyzz<F> xyzz_madd_signed(const Xyzz<F>& p, const typename F::E& x2, const typename F::E& y2, bool negate) {
    return xyzz_madd<F>(p, x2, negate ? F::neg(y2) : y2);
}

// G1: madd-2008-s with lazy reductions.  The negation is folded into a wide
// operand (4m - y2), R stays unreduced (< 6m) into its square and its product,
// X3 = R^2 - PPP - 2Q takes one normalization and one reduction, and
// Y3 = R (Q - X3) + Y (4m - PPP) is one double product with one Montgomery
// reduction.  Bounds: S2, U2, PP, PPP, Q, R^2 < 2m (products); P < 2m; R < 6m
// normalized; Q - X3 wide (< 6m); the double product < 36 m^2 + 8 m^2.  Outputs
// are normalized and < 2m, as from xyzz_madd.
template <>
HD Xyzz<FqOps> xyzz_madd_signed<FqOps>(const Xyzz<FqOps>& p, const Fq& x2, const Fq& y2, bool negate) {
    const Fq ys = negate ? sub_w(fp_zero<FqParams>(), y2) : y2;  // wide when negated, < 4m
    if (is_zero(p.ZZ)) return {x2, negate ? reduce(norm(ys)) : y2, fp_one<FqParams>(), fp_one<FqParams>()};
    const Fq U2 = mul(x2, p.ZZ);
    const Fq S2 = mul(ys, p.ZZZ);
    const Fq P = reduce(norm(sub_w(U2, p.X)));
    const Fq R = norm(sub_w(S2, p.Y));  // < 6m
    if (is_zero(P)) {
        if (is_zero(reduce(R))) return xyzz_mdbl<FqOps>(x2, negate ? reduce(norm(ys)) : y2);
        return xyzz_inf<FqOps>();
    }
    const Fq PP = sqr(P);
    const Fq PPP = mul(P, PP);
    const Fq Q = mul(p.X, PP);
    const Fq X3 = reduce(sub3_n(sqr(R), PPP, Q, Q));
    const Fq Y3 = mul2(sub_w(Q, X3), R, p.Y, norm(sub_w(fp_zero<FqParams>(), PPP)));
    return {X3, Y3, mul(p.ZZ, PP), mul(p.ZZZ, PPP)};
}

// P + Q (add-2008-s).
template <class F>
HD Xyzz<F> xyzz_add(const Xyzz<F>& p, const Xyzz<F>& q) {
    using E = typename F::E;
    if (F::is_zero(p.ZZ)) return q;
    if (F::is_zero(q.ZZ)) return p;
    const E U1 = F::mul(p.X, q.ZZ);
    const E U2 = F::mul(q.X, p.ZZ);
    const E S1 = F::mul(p.Y, q.ZZZ);
    const E S2 = F::mul(q.Y, p.ZZZ);
    const E P = F::sub(U2, U1);
    const E R = F::sub(S2, S1);
    if (F::is_zero(P)) {
        if (F::is_zero(R)) return xyzz_dbl<F>(p);
        return xyzz_inf<F>();
    }
    const E PP = F::sqr(P);
    const E PPP = F::mul(P, PP);
    const E Q = F::mul(U1, PP);
    const E X3 = F::sub(F::sub(F::sqr(R), PPP), F::add(Q, Q));
    const E Y3 = F::sub(F::mul(R, F::sub(Q, X3)), F::mul(S1, PPP));
    return {X3, Y3, F::mul(F::mul(p.ZZ, q.ZZ), PP), F::mul(F::mul(p.ZZZ, q.ZZZ), PPP)};
}

// G1: dbl-2008-s-1 with lazy reductions.  U = 2Y (< 4m) and M = 3X^2 (< 6m)
// stay unreduced into their products, X3 = M^2 - 2S takes one
// normalization and one reduction, Y3 = M (S - X3) + Y (4m - W) is one double
// product (< 36 m^2 + 8 m^2).  Inputs and outputs normalized and < 2m.
template <>
HD Xyzz<FqOps> xyzz_dbl<FqOps>(const Xyzz<FqOps>& p) {
    if (is_zero(p.ZZ) || is_zero(p.Y)) return xyzz_inf<FqOps>();
    const Fq U = norm(add_w(p.Y, p.Y));
    const Fq V = sqr(U);
    const Fq W = mul(U, V);
    const Fq S = mul(p.X, V);
    const Fq xx = sqr(p.X);
    const Fq M = norm(add_w(add_w(xx, xx), xx));
    const Fq X3 = reduce(sub2_n(sqr(M), S, S));
    const Fq Y3 = mul2(sub_w(S, X3), M, p.Y, norm(sub_w(fp_zero<FqParams>(), W)));
    return {X3, Y3, mul(V, p.ZZ), mul(W, p.ZZZ)};
}

// G1: add-2008-s with the same lazy steps as the mixed add below.
template <>
HD Xyzz<FqOps> xyzz_add<FqOps>(const Xyzz<FqOps>& p, const Xyzz<FqOps>& q) {
    if (is_zero(p.ZZ)) return q;
    if (is_zero(q.ZZ)) return p;
    const Fq U1 = mul(p.X, q.ZZ);
    const Fq U2 = mul(q.X, p.ZZ);
    const Fq S1 = mul(p.Y, q.ZZZ);
    const Fq S2 = mul(q.Y, p.ZZZ);
    const Fq P = reduce(norm(sub_w(U2, U1)));
    const Fq R = norm(sub_w(S2, S1));  // < 6m
    if (is_zero(P)) {
        if (is_zero(reduce(R))) return xyzz_dbl<FqOps>(p);
        return xyzz_inf<FqOps>();
    }
    const Fq PP = sqr(P);
    const Fq PPP = mul(P, PP);
    const Fq Q = mul(U1, PP);
    const Fq X3 = reduce(sub3_n(sqr(R), PPP, Q, Q));
    const Fq Y3 = mul2(sub_w(Q, X3), R, S1, norm(sub_w(fp_zero<FqParams>(), PPP)));
    return {X3, Y3, mul(mul(p.ZZ, q.ZZ), PP), mul(mul(p.ZZZ, q.ZZZ), PPP)};
}

template <class F>
HD bool xyzz_eq(const Xyzz<F>& p, const Xyzz<F>& q) {
    const bool pi = F::is_zero(p.ZZ), qi = F::is_zero(q.ZZ);
    if (pi || qi) return pi && qi;
    return F::is_zero(F::sub(F::mul(p.X, q.ZZ), F::mul(q.X, p.ZZ))) &&
           F::is_zero(F::sub(F::mul(p.Y, q.ZZZ), F::mul(q.Y, p.ZZZ)));
}

// [k]P for an unsigned integer given as 32-bit words, most significant bit
// first (plain double-and-add; used by validation and small scalings).
template <class F, int WORDS>
HD Xyzz<F> xyzz_mul_words(const Xyzz<F>& p, const uint32_t (&k)[WORDS], int bits) {
    Xyzz<F> r = xyzz_inf<F>();
    for (int b = bits - 1; b >= 0; --b) {
        r = xyzz_dbl<F>(r);
        if ((k[b >> 5] >> (b & 31)) & 1) r = xyzz_add<F>(r, p);
    }
    return r;
}

template <class F>
HD Xyzz<F> xyzz_mul_u64(const Xyzz<F>& p, uint64_t k) {
    const uint32_t w[2] = {uint32_t(k), uint32_t(k >> 32)};
    return xyzz_mul_words<F, 2>(p, w, 64);
}

// Affine (canonical, fully reduced) from XYZZ.
template <class F>
HD Affine<F> xyzz_to_affine(const Xyzz<F>& p) {
    using E = typename F::E;
    if (F::is_zero(p.ZZ)) return {F::zero(), F::zero(), true};
    // ZZZ^-1 = Z^-3;  Z^-1 = ZZZ^-1 * ZZ;  ZZ^-1 = (Z^-1)^2
    const E izzz = F::inv(p.ZZZ);
    const E iz = F::mul(izzz, p.ZZ);
    const E izz = F::sqr(iz);
    return {F::canon(F::mul(p.X, izz)), F::canon(F::mul(p.Y, izzz)), false};
}

template <class F>
HD bool affine_on_curve(const Affine<F>& a) {
    if (a.inf) return true;
    const auto lhs = F::sqr(a.y);
    const auto rhs = F::add(F::mul(F::sqr(a.x), a.x), F::b());
    return F::is_zero(F::sub(lhs, rhs));
}

// ---- G2 endomorphism psi (untwist-Frobenius-twist) --------------------------

HD Fq2 psi_cx() { return {fp_const<FqParams>(PSI_X_C0), fp_const<FqParams>(PSI_X_C1)}; }
HD Fq2 psi_cy() { return {fp_const<FqParams>(PSI_Y_C0), fp_const<FqParams>(PSI_Y_C1)}; }

HD Xyzz<Fq2Ops> psi(const Xyzz<Fq2Ops>& p) {
    return {f2_mul(f2_conj(p.X), psi_cx()), f2_mul(f2_conj(p.Y), psi_cy()), f2_conj(p.ZZ),
            f2_conj(p.ZZZ)};
}

// ---- memory helpers ------------------------------------------------------

// 64-byte G1 affine in memory: x || y as Montgomery values (< 2^256), plus
// the all-zero infinity convention (the canonical coordinates (0, 0) are not
// on the curve, so all-zero Montgomery words are never a finite point).
HD Affine<FqOps> g1_load_mont(const uint32_t* w) {
    uint32_t x[8], y[8];
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        x[i] = w[i];
        y[i] = w[8 + i];
        any |= x[i] | y[i];
    }
    return {unpack<FqParams>(x), unpack<FqParams>(y), any == 0};
}

HD void g1_store_mont(const Affine<FqOps>& a, uint32_t* w) {
    uint32_t x[8], y[8];
    if (a.inf) {
#pragma unroll
        for (int i = 0; i < 16; ++i) w[i] = 0;
        return;
    }
    pack(a.x, x);
    pack(a.y, y);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        w[i] = x[i];
        w[8 + i] = y[i];
    }
}

}  // namespace bn254
