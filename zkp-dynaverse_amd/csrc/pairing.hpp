// Optimal ate pairing on BN254 (host and device), for batched pairing-product
// checks: prod_i e(P_i, Q_i) == 1.
//
// Tower: Fq2 = Fq[u]/(u^2 + 1), Fq6 = Fq2[v]/(v^3 - xi), Fq12 = Fq6[w]/(w^2 - v),
// xi = 9 + u, so w^6 = xi and an Fq12 element c0 + c1 w (c_j = a0 + a1 v +
// a2 v^2) has the Fq2 coefficient (c_j)_i at w^(2i + j).
//
// Miller loop f_{6x+2,Q}(P) l_{T,pi(Q)}(P) l_{T',-pi^2(Q)}(P) over the
// binary digits of 6x + 2, with T in homogeneous projective coordinates on
// the twist E': y^2 = x^3 + 3/xi and the lines of the untwisted points
// (x w^2, y w^3) evaluated at P = (xp, yp).  With slope m the line is
// yp - (m xp) w + (m x_T - y_T) w^3; every line is scaled by a factor in Fq2,
// which the final exponentiation removes (Fq2 lies in Fq6, and a^(q^6-1) = 1
// there):
//   tangent at T = (X, Y, Z):  (2 Y Z^2) yp - (3 X^2 Z) xp w + (3 X^3 - 2 Y^2 Z) w^3
//   chord T, Q = (x2, y2):     (x2 Z - X) yp - (y2 Z - Y) xp w + ((y2 Z - Y) x2 - (x2 Z - X) y2) w^3
// Point steps: dbl-2007-bl and madd-1998-cmo (EFD, homogeneous, a = 0).
// Final exponentiation: f^((q^6 - 1)(q^2 + 1)) by conjugation, inversion and
// the q^2-power Frobenius, then the hard part (q^4 - q^2 + 1)/r through
// three exponentiations by x and the q-power Frobenius.  Several pairs share
// one Miller loop (one squaring chain) and one final exponentiation.
#pragma once
#include "curve.hpp"

// The Fq12-level and Miller-step functions stay out of line on the device:
// inlined, one check would be a single body of several hundred thousand
// instructions (minutes of compile time, the whole register file spilled).
// Fq6 products are inlined into them (10% faster than calling them).
#if defined(__HIPCC__)
#define HD_CALL __host__ __device__ inline __attribute__((noinline))
#else
#define HD_CALL inline
#endif

namespace bn254 {

struct Fq6 {
    Fq2 a0, a1, a2;
};
struct Fq12 {
    Fq6 c0, c1;
};

HD Fq2 f2_mul_xi(const Fq2& a) {  // (a0 + a1 u)(9 + u) = (9 a0 - a1) + (a0 + 9 a1) u
    const Fq a0x9 = add(dbl(dbl(dbl(a.c0))), a.c0), a1x9 = add(dbl(dbl(dbl(a.c1))), a.c1);
    return {sub(a0x9, a.c1), add(a.c0, a1x9)};
}

HD Fq6 f6_zero() { return {f2_zero(), f2_zero(), f2_zero()}; }
HD Fq6 f6_one() { return {f2_one(), f2_zero(), f2_zero()}; }
HD Fq6 f6_add(const Fq6& a, const Fq6& b) { return {f2_add(a.a0, b.a0), f2_add(a.a1, b.a1), f2_add(a.a2, b.a2)}; }
HD Fq6 f6_sub(const Fq6& a, const Fq6& b) { return {f2_sub(a.a0, b.a0), f2_sub(a.a1, b.a1), f2_sub(a.a2, b.a2)}; }
HD Fq6 f6_neg(const Fq6& a) { return {f2_neg(a.a0), f2_neg(a.a1), f2_neg(a.a2)}; }
HD Fq6 f6_mul_v(const Fq6& a) { return {f2_mul_xi(a.a2), a.a0, a.a1}; }  // a v

HD Fq6 f6_mul(const Fq6& a, const Fq6& b) {
    // Karatsuba over the three coefficients (6 Fq2 products).
    const Fq2 t0 = f2_mul(a.a0, b.a0), t1 = f2_mul(a.a1, b.a1), t2 = f2_mul(a.a2, b.a2);
    const Fq2 m12 = f2_sub(f2_sub(f2_mul(f2_add(a.a1, a.a2), f2_add(b.a1, b.a2)), t1), t2);
    const Fq2 m01 = f2_sub(f2_sub(f2_mul(f2_add(a.a0, a.a1), f2_add(b.a0, b.a1)), t0), t1);
    const Fq2 m02 = f2_sub(f2_sub(f2_mul(f2_add(a.a0, a.a2), f2_add(b.a0, b.a2)), t0), t2);
    return {f2_add(t0, f2_mul_xi(m12)), f2_add(m01, f2_mul_xi(t2)), f2_add(m02, t1)};
}

HD Fq6 f6_mul_f2(const Fq6& a, const Fq2& b) { return {f2_mul(a.a0, b), f2_mul(a.a1, b), f2_mul(a.a2, b)}; }

HD_CALL Fq6 f6_inv(const Fq6& a) {
    const Fq2 t0 = f2_sub(f2_sqr(a.a0), f2_mul_xi(f2_mul(a.a1, a.a2)));
    const Fq2 t1 = f2_sub(f2_mul_xi(f2_sqr(a.a2)), f2_mul(a.a0, a.a1));
    const Fq2 t2 = f2_sub(f2_sqr(a.a1), f2_mul(a.a0, a.a2));
    const Fq2 den = f2_add(f2_mul(a.a0, t0), f2_mul_xi(f2_add(f2_mul(a.a2, t1), f2_mul(a.a1, t2))));
    const Fq2 di = f2_inv(den);
    return {f2_mul(t0, di), f2_mul(t1, di), f2_mul(t2, di)};
}

HD Fq12 f12_one() { return {f6_one(), f6_zero()}; }

HD_CALL Fq12 f12_mul(const Fq12& a, const Fq12& b) {
    const Fq6 t0 = f6_mul(a.c0, b.c0), t1 = f6_mul(a.c1, b.c1);
    const Fq6 m = f6_sub(f6_sub(f6_mul(f6_add(a.c0, a.c1), f6_add(b.c0, b.c1)), t0), t1);
    return {f6_add(t0, f6_mul_v(t1)), m};
}

HD_CALL Fq12 f12_sqr(const Fq12& a) {
    // (c0 + c1 w)^2 = c0^2 + v c1^2 + 2 c0 c1 w, as (c0 + c1)(c0 + v c1) - ...
    const Fq6 p = f6_mul(a.c0, a.c1);
    const Fq6 s = f6_mul(f6_add(a.c0, a.c1), f6_add(a.c0, f6_mul_v(a.c1)));
    return {f6_sub(f6_sub(s, p), f6_mul_v(p)), f6_add(p, p)};
}

HD Fq12 f12_conj(const Fq12& a) { return {a.c0, f6_neg(a.c1)}; }

// Square of an element of the cyclotomic subgroup (a^(q^6 + 1) = 1)
// (Granger, Scott, "Faster squaring in the cyclotomic subgroup of sixth
// degree extensions", PKC 2010, sec. 3.2).  With y = w^3 (y^2 = xi), the
// element splits into three Fq4 = Fq2[y] parts A = z0 + z1 y, B = z2 + z3 y,
// C = z4 + z5 y on the coefficients z0 = c0.a0, z4 = c0.a1, z3 = c0.a2,
// z2 = c1.a0, z1 = c1.a1, z5 = c1.a2; the square needs only A^2, B^2, C^2
// (each u^2 + xi v^2 and 2uv from one Karatsuba product):
//   z0' = 3 (A^2)_0 - 2 z0,   z1' = 3 (A^2)_1 + 2 z1,
//   z2' = 3 xi (C^2)_1 + 2 z2, z3' = 3 (C^2)_0 - 2 z3,
//   z4' = 3 (B^2)_0 - 2 z4,   z5' = 3 (B^2)_1 + 2 z5.
HD void fq4_sqr(const Fq2& u, const Fq2& v, Fq2& r0, Fq2& r1) {  // (u + v y)^2
    const Fq2 uv = f2_mul(u, v);
    r0 = f2_sub(f2_sub(f2_mul(f2_add(u, v), f2_add(u, f2_mul_xi(v))), uv), f2_mul_xi(uv));
    r1 = f2_dbl(uv);
}

HD_CALL Fq12 f12_cyclotomic_sqr(const Fq12& a) {
    const Fq2 &z0 = a.c0.a0, &z4 = a.c0.a1, &z3 = a.c0.a2, &z2 = a.c1.a0, &z1 = a.c1.a1, &z5 = a.c1.a2;
    Fq2 a0, a1, b0, b1, c0, c1;
    fq4_sqr(z0, z1, a0, a1);
    fq4_sqr(z2, z3, b0, b1);
    fq4_sqr(z4, z5, c0, c1);
    const auto three_minus_two = [](const Fq2& t, const Fq2& z) {  // 3t - 2z
        const Fq2 d = f2_sub(t, z);
        return f2_add(f2_dbl(d), t);
    };
    const auto three_plus_two = [](const Fq2& t, const Fq2& z) {  // 3t + 2z
        const Fq2 d = f2_add(t, z);
        return f2_add(f2_dbl(d), t);
    };
    Fq12 r;
    r.c0.a0 = three_minus_two(a0, z0);
    r.c1.a1 = three_plus_two(a1, z1);
    r.c1.a0 = three_plus_two(f2_mul_xi(c1), z2);
    r.c0.a2 = three_minus_two(c0, z3);
    r.c0.a1 = three_minus_two(b0, z4);
    r.c1.a2 = three_plus_two(b1, z5);
    return r;
}

HD_CALL Fq12 f12_inv(const Fq12& a) {
    const Fq6 den = f6_sub(f6_mul(a.c0, a.c0), f6_mul_v(f6_mul(a.c1, a.c1)));
    const Fq6 di = f6_inv(den);
    return {f6_mul(a.c0, di), f6_neg(f6_mul(a.c1, di))};
}

// a^q: conjugate every Fq2 coefficient and scale the coefficient of w^k by
// xi^(k(q-1)/6).
HD Fq2 frob_c(int k) {
    switch (k) {
        case 1: return {fp_const<FqParams>(FROB_C1_0), fp_const<FqParams>(FROB_C1_1)};
        case 2: return {fp_const<FqParams>(FROB_C2_0), fp_const<FqParams>(FROB_C2_1)};
        case 3: return {fp_const<FqParams>(FROB_C3_0), fp_const<FqParams>(FROB_C3_1)};
        case 4: return {fp_const<FqParams>(FROB_C4_0), fp_const<FqParams>(FROB_C4_1)};
        default: return {fp_const<FqParams>(FROB_C5_0), fp_const<FqParams>(FROB_C5_1)};
    }
}

HD_CALL Fq12 f12_frobenius(const Fq12& a) {
    // w^(2i + j): c_j's coefficient i
    return {{f2_conj(a.c0.a0), f2_mul(f2_conj(a.c0.a1), frob_c(2)), f2_mul(f2_conj(a.c0.a2), frob_c(4))},
            {f2_mul(f2_conj(a.c1.a0), frob_c(1)), f2_mul(f2_conj(a.c1.a1), frob_c(3)),
             f2_mul(f2_conj(a.c1.a2), frob_c(5))}};
}

HD bool f12_is_one(const Fq12& a) {
    return eq(a.c0.a0.c0, fp_one<FqParams>()) && is_zero(a.c0.a0.c1) && f2_is_zero(a.c0.a1) &&
           f2_is_zero(a.c0.a2) && f2_is_zero(a.c1.a0) && f2_is_zero(a.c1.a1) && f2_is_zero(a.c1.a2);
}

struct G2Proj {
    Fq2 X, Y, Z;
};

// A line l0 + l1 w + l3 w^3 (Fq2 coefficients; w^1 and w^3 are c1's first
// two coefficients in the tower).
struct Line {
    Fq2 l0, l1, l3;
};

// f * line: (f0 + f1 w)(c0 + c1 w) with c0 = l0, c1 = l1 + l3 v; 15 Fq2 products.
HD_CALL Fq12 f12_mul_line(const Fq12& f, const Line& l) {
    const Fq6 t0 = f6_mul_f2(f.c0, l.l0);
    // f1 (l1 + l3 v) = (b0 l1 + xi b2 l3) + (b0 l3 + b1 l1) v + (b1 l3 + b2 l1) v^2
    const Fq6& b = f.c1;
    const Fq6 t1 = {f2_add(f2_mul(b.a0, l.l1), f2_mul_xi(f2_mul(b.a2, l.l3))),
                    f2_add(f2_mul(b.a0, l.l3), f2_mul(b.a1, l.l1)), f2_add(f2_mul(b.a1, l.l3), f2_mul(b.a2, l.l1))};
    // (f0 + f1)((l0 + l1) + l3 v)
    const Fq6 s = f6_add(f.c0, f.c1);
    const Fq2 m0 = f2_add(l.l0, l.l1);
    const Fq6 t2 = {f2_add(f2_mul(s.a0, m0), f2_mul_xi(f2_mul(s.a2, l.l3))),
                    f2_add(f2_mul(s.a0, l.l3), f2_mul(s.a1, m0)), f2_add(f2_mul(s.a1, l.l3), f2_mul(s.a2, m0))};
    return {f6_add(t0, f6_mul_v(t1)), f6_sub(f6_sub(t2, t0), t1)};
}

// Tangent line at T with its P-dependent factors left out (l0 stands for the
// coefficient of yp, l1 for that of xp), and T <- 2T.
HD_CALL Line miller_dbl_coeffs(G2Proj& t) {
    const Fq2 xx = f2_sqr(t.X), yy = f2_sqr(t.Y);
    const Fq2 yz = f2_mul(t.Y, t.Z);
    const Fq2 three_xx = f2_add(f2_dbl(xx), xx);
    // line: (2 Y Z^2) yp - (3 X^2 Z) xp w + (3 X^3 - 2 Y^2 Z) w^3
    const Line l = {f2_dbl(f2_mul(yz, t.Z)), f2_neg(f2_mul(three_xx, t.Z)),
                    f2_sub(f2_mul(three_xx, t.X), f2_dbl(f2_mul(yy, t.Z)))};
    // dbl-2007-bl, a = 0: w = 3XX, s = 2YZ, R = Y s, B = (X + R)^2 - XX - RR,
    // h = w^2 - 2B; X3 = h s, Y3 = w (B - h) - 2 RR, Z3 = s^3.
    const Fq2 s = f2_dbl(yz);
    const Fq2 ss = f2_sqr(s);
    const Fq2 r = f2_mul(t.Y, s);
    const Fq2 rr = f2_sqr(r);
    const Fq2 b = f2_sub(f2_sub(f2_sqr(f2_add(t.X, r)), xx), rr);
    const Fq2 h = f2_sub(f2_sqr(three_xx), f2_dbl(b));
    t = {f2_mul(h, s), f2_sub(f2_mul(three_xx, f2_sub(b, h)), f2_dbl(rr)), f2_mul(s, ss)};
    return l;
}

// Chord through T and the affine Q = (x2, y2) in the same form, and T <- T + Q.
HD_CALL Line miller_add_coeffs(G2Proj& t, const Fq2& x2, const Fq2& y2) {
    const Fq2 theta = f2_sub(f2_mul(y2, t.Z), t.Y);  // y2 Z - Y
    const Fq2 lambda = f2_sub(f2_mul(x2, t.Z), t.X); // x2 Z - X
    const Line l = {lambda, f2_neg(theta), f2_sub(f2_mul(theta, x2), f2_mul(lambda, y2))};
    // madd-1998-cmo: uu = u^2, vv = v^2, vvv = v vv, R = vv X,
    // A = uu Z - vvv - 2R; X3 = v A, Y3 = u (R - A) - vvv Y, Z3 = vvv Z.
    const Fq2 uu = f2_sqr(theta), vv = f2_sqr(lambda);
    const Fq2 vvv = f2_mul(lambda, vv);
    const Fq2 r = f2_mul(vv, t.X);
    const Fq2 a = f2_sub(f2_sub(f2_mul(uu, t.Z), vvv), f2_dbl(r));
    t = {f2_mul(lambda, a), f2_sub(f2_mul(theta, f2_sub(r, a)), f2_mul(vvv, t.Y)), f2_mul(vvv, t.Z)};
    return l;
}

// The line at P = (xp, yp).
HD Line line_at(const Line& c, const Fq& xp, const Fq& yp) { return {f2_mul_fq(c.l0, yp), f2_mul_fq(c.l1, xp), c.l3}; }

HD Line miller_dbl(G2Proj& t, const Fq& xp, const Fq& yp) { return line_at(miller_dbl_coeffs(t), xp, yp); }
HD Line miller_add(G2Proj& t, const Fq2& x2, const Fq2& y2, const Fq& xp, const Fq& yp) {
    return line_at(miller_add_coeffs(t, x2, y2), xp, yp);
}

// pi(Q) on the twist: (conj(x) xi^((q-1)/3), conj(y) xi^((q-1)/2)).
HD void twist_frobenius(const Fq2& x, const Fq2& y, Fq2& ox, Fq2& oy) {
    ox = f2_mul(f2_conj(x), frob_c(2));
    oy = f2_mul(f2_conj(y), frob_c(3));
}

struct PairingInput {
    Fq xp, yp;   // P, affine, finite
    Fq2 x2, y2;  // Q on the twist, affine, finite
};

constexpr int PAIRING_MAX_TERMS = 16;

constexpr int ate_popcount() {
    int c = 0;
    for (int i = 0; i < 64; ++i) c += int((PAIRING_ATE_LO >> i) & 1u);
    for (int i = 0; i < 32; ++i) c += int((PAIRING_ATE_HI >> i) & 1u);
    return c;
}
// Lines of one pair's Miller loop: one tangent per bit below the top, one
// chord per set bit below the top, the pi(Q) and -pi^2(Q) chords.
constexpr int PAIRING_LINES = (PAIRING_ATE_BITS - 1) + (ate_popcount() - 1) + 2;

HD uint32_t ate_bit(int i) {
    return i >= 64 ? (PAIRING_ATE_HI >> (i - 64)) & 1u : uint32_t(PAIRING_ATE_LO >> i) & 1u;
}

// The P-independent line coefficients of Q's Miller loop, in loop order
// (PAIRING_LINES entries): for a Q shared by many pairings.
HD_CALL void prepare_lines(const Fq2& x2, const Fq2& y2, Line* out) {
    G2Proj t = {x2, y2, f2_one()};
    int n = 0;
    for (int i = PAIRING_ATE_BITS - 2; i >= 0; --i) {
        out[n++] = miller_dbl_coeffs(t);
        if (ate_bit(i)) out[n++] = miller_add_coeffs(t, x2, y2);
    }
    Fq2 q1x, q1y, q2x, q2y;
    twist_frobenius(x2, y2, q1x, q1y);
    twist_frobenius(q1x, q1y, q2x, q2y);
    out[n++] = miller_add_coeffs(t, q1x, q1y);
    out[n++] = miller_add_coeffs(t, q2x, f2_neg(q2y));
}

// prod_j f_j for k finite pairs, sharing the squaring of the accumulator
// (one Miller loop over all pairs).  `prepared` (optional): one more pair
// whose Q's lines come from prepare_lines, evaluated at (pxp, pyp).
template <class LineSource>
HD Fq12 miller_product_impl(const PairingInput* in, int k, const LineSource& prepared, bool has_prepared,
                            const Fq& pxp, const Fq& pyp) {
    G2Proj t[PAIRING_MAX_TERMS];
    for (int j = 0; j < k; ++j) t[j] = {in[j].x2, in[j].y2, f2_one()};
    Fq12 f = f12_one();
    int n = 0;
    for (int i = PAIRING_ATE_BITS - 2; i >= 0; --i) {
        if (i != PAIRING_ATE_BITS - 2) f = f12_sqr(f);
        if (has_prepared) f = f12_mul_line(f, line_at(prepared(n++), pxp, pyp));
        for (int j = 0; j < k; ++j) f = f12_mul_line(f, miller_dbl(t[j], in[j].xp, in[j].yp));
        if (ate_bit(i)) {
            if (has_prepared) f = f12_mul_line(f, line_at(prepared(n++), pxp, pyp));
            for (int j = 0; j < k; ++j) f = f12_mul_line(f, miller_add(t[j], in[j].x2, in[j].y2, in[j].xp, in[j].yp));
        }
    }
    if (has_prepared) {
        f = f12_mul_line(f, line_at(prepared(n++), pxp, pyp));
        f = f12_mul_line(f, line_at(prepared(n++), pxp, pyp));
    }
    for (int j = 0; j < k; ++j) {
        Fq2 q1x, q1y, q2x, q2y;
        twist_frobenius(in[j].x2, in[j].y2, q1x, q1y);
        twist_frobenius(q1x, q1y, q2x, q2y);
        f = f12_mul_line(f, miller_add(t[j], q1x, q1y, in[j].xp, in[j].yp));
        f = f12_mul_line(f, miller_add(t[j], q2x, f2_neg(q2y), in[j].xp, in[j].yp));
    }
    return f;
}

struct NoLines {
    HD Line operator()(int) const { return {}; }
};

HD_CALL Fq12 miller_product(const PairingInput* in, int k) {
    return miller_product_impl(in, k, NoLines{}, false, fp_zero<FqParams>(), fp_zero<FqParams>());
}

// Miller value of one pair.
HD_CALL Fq12 miller_loop(const Fq& xp, const Fq& yp, const Fq2& x2, const Fq2& y2) {
    const PairingInput in = {xp, yp, x2, y2};
    return miller_product(&in, 1);
}

// f^x for the BN parameter x (63 bits), square-and-multiply, f in the
// cyclotomic subgroup.
HD_CALL Fq12 f12_pow_x(const Fq12& f) {
    Fq12 r = f;
    for (int i = 61; i >= 0; --i) {
        r = f12_cyclotomic_sqr(r);
        if ((BN_SEED >> i) & 1u) r = f12_mul(r, f);
    }
    return r;
}

// f^((q^12 - 1)/r).  Easy part f^((q^6 - 1)(q^2 + 1)); the result a lies in
// the cyclotomic subgroup, where a^-1 = conj(a).  Hard part
// (q^4 - q^2 + 1)/r = l0 + l1 q + l2 q^2 + l3 q^3 with
//   l3 = 1, l2 = 6x^2 + 1, l1 = -36x^3 - 18x^2 - 12x + 1,
//   l0 = -36x^3 - 30x^2 - 18x - 2
// (Scott, Benger, Charlemagne, Dominguez Perez, Kachisa, "On the final
// exponentiation for calculating pairings on ordinary elliptic curves",
// 2009), from a^x, a^(x^2), a^(x^3) and the q-power Frobenius.
HD_CALL Fq12 final_exponentiation(const Fq12& f) {
    Fq12 a = f12_mul(f12_conj(f), f12_inv(f));
    a = f12_mul(f12_frobenius(f12_frobenius(a)), a);
    const Fq12 x1 = f12_pow_x(a), x2 = f12_pow_x(x1), x3 = f12_pow_x(x2);
    const Fq12 x3_4 = f12_cyclotomic_sqr(f12_cyclotomic_sqr(x3));
    const Fq12 x3_36 = f12_mul(f12_cyclotomic_sqr(f12_cyclotomic_sqr(f12_cyclotomic_sqr(x3_4))), x3_4);
    const Fq12 x2_2 = f12_cyclotomic_sqr(x2);
    const Fq12 x2_6 = f12_mul(f12_cyclotomic_sqr(x2_2), x2_2);
    const Fq12 x2_12 = f12_cyclotomic_sqr(x2_6);
    const Fq12 x2_18 = f12_mul(x2_12, x2_6);
    const Fq12 x2_30 = f12_mul(x2_18, x2_12);
    const Fq12 x1_2 = f12_cyclotomic_sqr(x1);
    const Fq12 x1_6 = f12_mul(f12_cyclotomic_sqr(x1_2), x1_2);
    const Fq12 x1_12 = f12_cyclotomic_sqr(x1_6);
    const Fq12 x1_18 = f12_mul(x1_12, x1_6);
    const Fq12 p0 = f12_conj(f12_mul(f12_mul(x3_36, x2_30), f12_mul(x1_18, f12_cyclotomic_sqr(a))));
    const Fq12 p1 = f12_mul(f12_conj(f12_mul(f12_mul(x3_36, x2_18), x1_12)), a);
    const Fq12 p2 = f12_mul(x2_6, a);
    const Fq12 q2 = f12_frobenius(f12_frobenius(p2));
    const Fq12 q3 = f12_frobenius(f12_frobenius(f12_frobenius(a)));
    return f12_mul(f12_mul(p0, f12_frobenius(p1)), f12_mul(q2, q3));
}

}  // namespace bn254
