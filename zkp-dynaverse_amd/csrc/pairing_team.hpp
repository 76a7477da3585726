// Pairing arithmetic spread over one wavefront ("team"): one pairing-product
// check per wave of 64 lanes, with the Fq12 values in LDS.
//
// A lane that runs a whole check alone issues every Fq product of it in
// sequence (about 25 000 products, 27 ms on MI355X whether 1 or 4096 checks
// run).  Here every Fq12 product is one round of a team of one wave: lane
// 6i + j (< 36) forms the Fq2 product a_i b_j, then lane 2k + e (< 12) sums
// component e of the coefficient of w^k.  A check then costs about 630 such
// rounds.
//
// Representation: W12 holds an Fq12 element in the basis 1, w, .., w^5 over
// Fq2 (w^6 = xi = 9 + u), so a product is a plain polynomial product folded
// by w^6 = xi.  The tower form of pairing.hpp (c0 + c1 w, c_j in Fq6 over
// v = w^2) has the coefficient of w^(2i + j) at c_j.a_i.
//
// Every operation is run by all lanes of a one-wave block (they synchronise
// with __syncthreads, which for one wave only orders the LDS accesses).  Operands are LDS slots; an output may be an input (inputs
// are read before the first barrier, outputs written after).
#pragma once
#include "g1mem.hpp"
#include "pairing.hpp"

namespace bn254 {

struct W12 {
    Fq2 c[6];
};

// The team's LDS: W12 slots, the Fq2 products of the current round, the line
// scale factors and a few flags.  Slot numbers below.
constexpr int TW_F = 0, TW_G = 1, TW_H = 2, TW_OUT = 3, TW_A = 4, TW_X1 = 5, TW_X2 = 6, TW_X3 = 7, TW_T1 = 8,
              TW_T2 = 9, TW_T3 = 10, TW_T4 = 11, TW_T5 = 12, TW_T6 = 13, TW_T7 = 14, TW_U = 15, TW_SLOTS = 16;

// Threads of one team: one wave.
constexpr int TEAM_THREADS = 64;

struct TeamShared {
    Fq2 prod[36];
    W12 w[TW_SLOTS];
    W12 line[PAIRING_MAX_TERMS];  // pair p's current line, evaluated (w^2, w^4, w^5 stay zero)
    Fq2 scale[PAIRING_MAX_TERMS][3];  // pair p's line scale factors (s, 0) for l0, l1, l3
    const uint32_t* lines[PAIRING_MAX_TERMS];  // pair p's prepared lines
    uint32_t live[PAIRING_MAX_TERMS];          // pair p contributes (both points finite)
    uint32_t flags[4];  // kernel-specific
};

// One team per block: the block's LDS.
__shared__ TeamShared team_lds;

__device__ __forceinline__ void team_sync() { __syncthreads(); }

__device__ __forceinline__ Fq nine_times(const Fq& x) { return add(dbl(dbl(dbl(x))), x); }

__device__ __forceinline__ Fq12 w12_to_tower(const W12& a) {
    return {{a.c[0], a.c[2], a.c[4]}, {a.c[1], a.c[3], a.c[5]}};
}

__device__ __forceinline__ W12 tower_to_w12(const Fq12& a) {
    return {{a.c0.a0, a.c1.a0, a.c0.a1, a.c1.a1, a.c0.a2, a.c1.a2}};
}

// a^-1 for one lane, inline (f12_inv is an out-of-line call, whose register
// saves a kernel pays around every call site): a^-1 = conj(a) / N with
// N = a conj(a) = c0^2 - v c1^2 in Fq6, and N^-1 by the adjugate over Fq2.
__device__ __forceinline__ W12 w12_inv(const W12& a) {
    const Fq6 c0 = {a.c[0], a.c[2], a.c[4]}, c1 = {a.c[1], a.c[3], a.c[5]};
    const Fq6 n = f6_sub(f6_mul(c0, c0), f6_mul_v(f6_mul(c1, c1)));
    const Fq2 t0 = f2_sub(f2_sqr(n.a0), f2_mul_xi(f2_mul(n.a1, n.a2)));
    const Fq2 t1 = f2_sub(f2_mul_xi(f2_sqr(n.a2)), f2_mul(n.a0, n.a1));
    const Fq2 t2 = f2_sub(f2_sqr(n.a1), f2_mul(n.a0, n.a2));
    const Fq2 den = f2_add(f2_mul(n.a0, t0), f2_mul_xi(f2_add(f2_mul(n.a2, t1), f2_mul(n.a1, t2))));
    const Fq2 di = f2_inv(den);
    const Fq6 ni = {f2_mul(t0, di), f2_mul(t1, di), f2_mul(t2, di)};
    const Fq6 r0 = f6_mul(c0, ni), r1 = f6_neg(f6_mul(c1, ni));
    return {{r0.a0, r1.a0, r0.a1, r1.a1, r0.a2, r1.a2}};
}

// ---- the team's operations --------------------------------------------------
//
// A check is a program of operations on the W12 slots (team_program below),
// run by one loop with every operation's body inlined once, so the wave
// never makes a function call (a call would save and restore the live
// registers through scratch memory around every product).

// out = a * b: lane 6i + j (< 36) forms a_i b_j, then lane 2k + e
// (< 12) forms component e of the coefficient of w^k.
__device__ __forceinline__ void team_mul(TeamShared& s, const W12& a, const W12& b, W12& out) {
    const int lane = threadIdx.x;
    if (lane < 36) s.prod[lane] = f2_mul(a.c[lane / 6], b.c[lane % 6]);
    team_sync();
    if (lane < 12) {
        const int k = lane >> 1, e = lane & 1;
        // coefficient k: the sum over i + j = k (component e) plus xi times
        // the sum over i + j = k + 6.  Products are normalized and < 2m, so
        // six of them add limb-wise in 32 bits (limbs < 6 * 2^29).
        Fq lo = fp_zero<FqParams>(), h0 = fp_zero<FqParams>(), h1 = fp_zero<FqParams>();
        for (int i = 0; i < 6; ++i) {
            if (i <= k) {
                const Fq2& p = s.prod[6 * i + (k - i)];
                lo = add_w(lo, e ? p.c1 : p.c0);
            } else {
                const Fq2& p = s.prod[6 * i + (k + 6 - i)];
                h0 = add_w(h0, p.c0);
                h1 = add_w(h1, p.c1);
            }
        }
        // xi (h0 + h1 u) = (9 h0 - h1) + (h0 + 9 h1) u.  Component e is
        // lo + 9 x +- y (x = h_e, y = h_{1-e}); e = 0 adds 16m against the
        // subtraction.  Limb-wise in 64 bits: |terms| < 2^35; the value lies in
        // [0, 12m + 90m + 10m + 16m) = [0, 128m) (five products at most in the
        // h sums, each < 2m), below 2^261, so one signed carry pass normalizes
        // it and one reduction brings it below 2m.
        const Fq x = e ? h1 : h0, y = e ? h0 : h1;
        Fq r;
        int64_t carry = 0;
#pragma unroll
        for (int q = 0; q < LIMBS; ++q) {
            int64_t v = int64_t(lo.l[q]) + 9 * int64_t(x.l[q]) + carry;
            v += e ? int64_t(y.l[q]) : int64_t(16) * FqParams::MOD[q] - int64_t(y.l[q]);
            if (q < LIMBS - 1) {
                r.l[q] = uint32_t(v) & LIMB_MASK;
                carry = v >> LIMB_BITS;  // arithmetic shift
            } else {
                r.l[q] = uint32_t(v);
            }
        }
        r = reduce(r);
        if (e) out.c[k].c1 = r;
        else out.c[k].c0 = r;
    }
    team_sync();
}

enum TeamOp : uint32_t {
    TOP_MUL = 0,     // w[out] = w[a] * w[b]
    TOP_COPY = 1,    // w[out] = w[a]
    TOP_ONE = 2,     // w[out] = 1
    TOP_CONJ = 3,    // w[out] = w[a]^(q^6)
    TOP_FROB = 4,    // w[out] = w[a]^q
    TOP_INV = 5,     // w[out] = w[a]^-1
    TOP_LINES = 6,   // line[p] = line n of every live pair p < a, evaluated
    TOP_MUL_LINE = 7,  // w[TW_F] *= line[a] when pair a is live
};

__host__ __device__ constexpr uint32_t team_op(uint32_t op, uint32_t a, uint32_t b, uint32_t out, uint32_t n = 0) {
    return op | (a << 4) | (b << 9) | (out << 14) | (n << 19);
}

// Before a program: every pair's line has zero coefficients at w^2, w^4
// and w^5 (prepared lines are l0 + l1 w + l3 w^3).
__device__ __forceinline__ void team_clear_lines(TeamShared& s) {
    const int lane = threadIdx.x;
    if (lane < 3 * PAIRING_MAX_TERMS) s.line[lane / 3].c[lane % 3 == 0 ? 2 : lane % 3 + 3] = f2_zero();
    team_sync();
}

// Line n of pairs 0 .. k-1 (prepared lines in the layout of prepare_lines,
// PAIRING_LINES x {l0, l1, l3}, 48 Montgomery words per line), pair p's
// coefficients scaled by scale[p][0..2]; lane 3p + q forms coefficient q.
__device__ __forceinline__ void team_eval_lines(TeamShared& s, int k, int n) {
    const int lane = threadIdx.x;
    if (lane < 3 * k) {
        const int p = lane / 3, q = lane % 3;
        if (s.live[p]) {
            const uint32_t* src = s.lines[p] + n * 48 + q * 16;
            const Fq2 c = {dvgpu::ldfq(src), dvgpu::ldfq(src + 8)};
            s.line[p].c[q == 2 ? 3 : q] = f2_mul(c, s.scale[p][q]);  // w^0, w^1, w^3
        }
    }
    team_sync();
}

// Runs `len` operations of `prog` (pairs, lines and scales set in s).
__device__ __forceinline__ void team_run(TeamShared& s, const uint32_t* __restrict__ prog, int len) {
    const int lane = threadIdx.x;
    for (int pc = 0; pc < len; ++pc) {
        const uint32_t op = prog[pc];
        const int kind = int(op & 15u), a = int((op >> 4) & 31u), b = int((op >> 9) & 31u),
                  out = int((op >> 14) & 31u), n = int(op >> 19);
        switch (kind) {
            case TOP_MUL:
                team_mul(s, s.w[a], s.w[b], s.w[out]);
                break;
            case TOP_MUL_LINE:
                if (s.live[a]) team_mul(s, s.w[TW_F], s.line[a], s.w[TW_F]);
                break;
            case TOP_LINES:
                team_eval_lines(s, a, n);
                break;
            case TOP_INV: {
                if (lane == 0) s.w[out] = w12_inv(s.w[a]);
                team_sync();
                break;
            }
            default: {  // the coefficient-wise operations
                Fq2 v = f2_zero();
                if (lane < 6) {
                    const Fq2 x = s.w[a].c[lane];
                    if (kind == TOP_COPY) v = x;
                    else if (kind == TOP_ONE) v = lane ? f2_zero() : f2_one();
                    else if (kind == TOP_CONJ) v = (lane & 1) ? f2_neg(x) : x;
                    else v = lane ? f2_mul(f2_conj(x), frob_c(lane)) : f2_conj(x);  // TOP_FROB
                }
                team_sync();
                if (lane < 6) s.w[out].c[lane] = v;
                team_sync();
                break;
            }
        }
    }
}

// The program of one pairing-product check over k prepared pairs: the
// Miller product into w[TW_F] (the loop order of miller_product_impl: the
// accumulator squared once per bit below the top, each live pair's tangent,
// its chord where the bit is set, the two Frobenius chords), then the final
// exponentiation into w[TW_OUT] (the chain of final_exponentiation).
// Returns the number of operations (written to out when out is not null).
inline int team_program(int pairs, uint32_t* out) {
    int k = 0;
    auto op = [&](uint32_t kind, int a = 0, int b = 0, int o = 0, int n = 0) {
        if (out) out[k] = team_op(kind, uint32_t(a), uint32_t(b), uint32_t(o), uint32_t(n));
        ++k;
    };
    auto lines = [&](int n) {
        op(TOP_LINES, pairs, 0, 0, n);
        for (int p = 0; p < pairs; ++p) op(TOP_MUL_LINE, p);
    };
    auto mul = [&](int a, int b, int o) { op(TOP_MUL, a, b, o); };
    auto pow_x = [&](int src, int o) {
        op(TOP_COPY, src, 0, o);
        for (int i = 61; i >= 0; --i) {
            mul(o, o, o);
            if ((BN_SEED >> i) & 1u) mul(o, src, o);
        }
    };
    op(TOP_ONE, 0, 0, TW_F);
    int n = 0;
    for (int i = PAIRING_ATE_BITS - 2; i >= 0; --i) {
        if (i != PAIRING_ATE_BITS - 2) mul(TW_F, TW_F, TW_F);
        lines(n++);
        if (ate_bit(i)) lines(n++);
    }
    lines(n++);
    lines(n++);
    // easy part: a = f^((q^6 - 1)(q^2 + 1))
    op(TOP_INV, TW_F, 0, TW_U);
    op(TOP_CONJ, TW_F, 0, TW_A);
    mul(TW_A, TW_U, TW_A);
    op(TOP_FROB, TW_A, 0, TW_U);
    op(TOP_FROB, TW_U, 0, TW_U);
    mul(TW_U, TW_A, TW_A);
    // hard part
    pow_x(TW_A, TW_X1);
    pow_x(TW_X1, TW_X2);
    pow_x(TW_X2, TW_X3);
    mul(TW_X3, TW_X3, TW_T1);
    mul(TW_T1, TW_T1, TW_T1);  // x3^4
    mul(TW_T1, TW_T1, TW_T2);
    mul(TW_T2, TW_T2, TW_T2);
    mul(TW_T2, TW_T2, TW_T2);
    mul(TW_T2, TW_T1, TW_T2);  // x3^36
    mul(TW_X2, TW_X2, TW_T3);  // x2^2
    mul(TW_T3, TW_T3, TW_T4);
    mul(TW_T4, TW_T3, TW_T4);  // x2^6
    mul(TW_T4, TW_T4, TW_T5);  // x2^12
    mul(TW_T5, TW_T4, TW_T6);  // x2^18
    mul(TW_T6, TW_T5, TW_T7);  // x2^30
    mul(TW_X1, TW_X1, TW_T1);  // x1^2
    mul(TW_T1, TW_T1, TW_T3);
    mul(TW_T3, TW_T1, TW_T3);  // x1^6
    mul(TW_T3, TW_T3, TW_T5);  // x1^12
    mul(TW_T5, TW_T3, TW_X3);  // x1^18
    // p0 = conj(x3^36 x2^30 x1^18 a^2)
    mul(TW_A, TW_A, TW_U);
    mul(TW_X3, TW_U, TW_U);
    mul(TW_T2, TW_T7, TW_T1);
    mul(TW_T1, TW_U, TW_T1);
    op(TOP_CONJ, TW_T1, 0, TW_T1);
    // p1 = conj(x3^36 x2^18 x1^12) a
    mul(TW_T2, TW_T6, TW_U);
    mul(TW_U, TW_T5, TW_U);
    op(TOP_CONJ, TW_U, 0, TW_U);
    mul(TW_U, TW_A, TW_U);
    // q2 = (x2^6 a)^(q^2), q3 = a^(q^3)
    mul(TW_T4, TW_A, TW_T3);
    op(TOP_FROB, TW_T3, 0, TW_T3);
    op(TOP_FROB, TW_T3, 0, TW_T3);
    op(TOP_FROB, TW_A, 0, TW_T4);
    op(TOP_FROB, TW_T4, 0, TW_T4);
    op(TOP_FROB, TW_T4, 0, TW_T4);
    // p0 p1^q q2 q3
    op(TOP_FROB, TW_U, 0, TW_U);
    mul(TW_T1, TW_U, TW_T1);
    mul(TW_T3, TW_T4, TW_T3);
    mul(TW_T1, TW_T3, TW_OUT);
    return k;
}

}  // namespace bn254
