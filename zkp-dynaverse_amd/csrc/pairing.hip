// Batched pairing-product checks (include/dynaverse_pairing.h): one thread per
// group of pairs; Miller loops multiplied into one Fq12, one final
// exponentiation, the product compared with one.
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dynaverse_pairing.h"
#include "g1mem.hpp"
#include "glv.hpp"
#include "owner.hpp"
#include "pairing.hpp"
#include "pairing_team.hpp"

namespace dvgpu {

// Prepared G2 lines kept per owner: the lines of a G2 point depend on the
// point alone (arkworks' G2Prepared), and the verifiers' G2 points repeat
// call after call (the generator; a session's frozen divisor [tau - z] g2).
// Slots are handed out in order and all of them are dropped together when a
// call needs more than remain.
constexpr uint32_t LINE_SLOTS = 4096;

struct PairingCache {
    std::unordered_map<std::string, uint32_t> slot_of;  // 128-byte wire point -> slot
    DeviceBuffer lines;                                 // LINE_SLOTS x (lines of one point)
    std::string g1_key;                                 // 64-byte wire point of g1_table
    DeviceBuffer g1_table;                              // [2^i] g1 for i < 256, XYZZ words
    DeviceBuffer staged;                                // points to prepare, their slots
    DeviceBuffer program;  // team_program for 1..PAIRING_MAX_TERMS pairs, then (offset, length) per count
    uint32_t program_words = 0;
};

namespace {

using namespace bn254;

__device__ __forceinline__ Fq ld_canon_fq(const uint32_t* p) {
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) w[i] = p[i];
    return to_mont<FqParams>(w);
}

__device__ __forceinline__ void st_canon_fq(uint32_t* p, const Fq& v) {
    uint32_t w[8];
    from_mont(v, w);
    for (int i = 0; i < 8; ++i) p[i] = w[i];
}

__device__ __forceinline__ bool all_zero(const uint32_t* p, int n) {
    uint32_t any = 0;
    for (int i = 0; i < n; ++i) any |= p[i];
    return !any;
}

__device__ __forceinline__ bool canonical_fq(const uint32_t* p) {
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) w[i] = p[i];
    return is_canonical<FqParams>(w);
}

// Pair j at p (G1, 16 words) and q (G2, 32 words), canonical LE: false when
// a coordinate is not canonical or a point is off its curve.  `finite` says
// whether both points are finite (a pair with an identity contributes one).
__device__ __forceinline__ bool load_pair(const uint32_t* p, const uint32_t* q, PairingInput& e, bool& finite) {
    const bool p0 = all_zero(p, 16), q0 = all_zero(q, 32);
    bool pv = p0, qv = q0;
    if (!p0 && canonical_fq(p) && canonical_fq(p + 8)) {
        e.xp = ld_canon_fq(p);
        e.yp = ld_canon_fq(p + 8);
        pv = affine_on_curve<FqOps>({e.xp, e.yp, false});
    }
    if (!q0 && canonical_fq(q) && canonical_fq(q + 8) && canonical_fq(q + 16) && canonical_fq(q + 24)) {
        e.x2 = {ld_canon_fq(q), ld_canon_fq(q + 8)};
        e.y2 = {ld_canon_fq(q + 16), ld_canon_fq(q + 24)};
        qv = affine_on_curve<Fq2Ops>({e.x2, e.y2, false});
    }
    finite = !p0 && !q0;
    return pv && qv;
}

__global__ void __launch_bounds__(64) k_pairing_check(const uint32_t* __restrict__ g1, const uint32_t* __restrict__ g2,
                                                      uint32_t terms, uint64_t count, uint8_t* __restrict__ ok,
                                                      uint32_t* __restrict__ gt) {
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= count) return;
    PairingInput in[PAIRING_MAX_TERMS];
    int k = 0;
    bool valid = true;
    for (uint32_t j = 0; j < terms; ++j) {
        PairingInput e;
        bool finite;
        valid = load_pair(g1 + (g * terms + j) * 16, g2 + (g * terms + j) * 32, e, finite) && valid;
        if (finite) in[k++] = e;  // e(O, Q) = e(P, O) = 1
    }
    if (!valid) {  // reported to the host, which fails the call
        ok[g] = 2;
        return;
    }
    const Fq12 r = final_exponentiation(k ? miller_product(in, k) : f12_one());
    ok[g] = f12_is_one(r) ? 1 : 0;
    if (gt) {
        uint32_t* o = gt + g * 96;
        const Fq6* c[2] = {&r.c0, &r.c1};
        for (int k = 0; k < 2; ++k) {
            const Fq2* a[3] = {&c[k]->a0, &c[k]->a1, &c[k]->a2};
            for (int i = 0; i < 3; ++i) {
                st_canon_fq(o + (k * 3 + i) * 16, a[i]->c0);
                st_canon_fq(o + (k * 3 + i) * 16 + 8, a[i]->c1);
            }
        }
    }
}

// One thread per equation e(L) == prod_j e(R_j) (verification.rs:37-103),
// pairs [offsets[i], offsets[i + 1]) with the left pair first.
// Product mode (verification.rs:79-103): one Miller product over L and the
// right pairs with negated G1 points, one final exponentiation, compared with
// one.  Reference mode (verification.rs:46-53): every pair reduced on its
// own, the right-hand values multiplied in GT and compared with the left
// value (A == B as A * conj(B) == 1: reduced values are unitary).
// Thread t takes equation order[t] (equations ordered by pair count, so a
// wave's lanes run Miller products of one length).  ok[i] = 2 flags a bad
// point for the host.
__global__ void __launch_bounds__(64) k_pairing_equations(const uint32_t* __restrict__ g1,
                                                          const uint32_t* __restrict__ g2,
                                                          const uint32_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ order, uint64_t count,
                                                          int reference_mode, uint8_t* __restrict__ ok) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint64_t i = order[t];
    const uint32_t first = offsets[i], terms = offsets[i + 1] - first;
    PairingInput in[PAIRING_MAX_TERMS];
    bool live[PAIRING_MAX_TERMS];
    int k = 0;
    bool valid = true;
    for (uint32_t j = 0; j < terms; ++j) {
        PairingInput e;
        bool finite;
        valid = load_pair(g1 + size_t(first + j) * 16, g2 + size_t(first + j) * 32, e, finite) && valid;
        if (!reference_mode) {
            if (!finite) continue;
            if (j) e.yp = neg(e.yp);  // bilinearity moves a right-hand pair to the left
            in[k++] = e;
        } else {
            in[j] = e;
            live[j] = finite;
        }
    }
    if (!valid) {
        ok[i] = 2;
        return;
    }
    if (!reference_mode) {
        ok[i] = f12_is_one(final_exponentiation(k ? miller_product(in, k) : f12_one())) ? 1 : 0;
        return;
    }
    const Fq12 left = live[0] ? final_exponentiation(miller_product(&in[0], 1)) : f12_one();
    Fq12 right = f12_one();
    for (uint32_t j = 1; j < terms; ++j)
        if (live[j]) right = f12_mul(right, final_exponentiation(miller_product(&in[j], 1)));
    ok[i] = f12_is_one(f12_mul(left, f12_conj(right))) ? 1 : 0;
}

// ---- opening-cache check (rebuilding/machine.rs:357-382) --------------------

__device__ __forceinline__ bool canonical_fr(const uint32_t* p) {
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) w[i] = p[i];
    return is_canonical<FrParams>(w);
}

// Divisor D_i = [tau]g2 - [z_i] g2_0 as affine Montgomery words (x.c0, x.c1,
// y.c0, y.c1; all-zero words = infinity).  flags[i] = 2 for a non-canonical z_i.
__global__ void __launch_bounds__(64) k_cache_divisors(const uint32_t* __restrict__ z, uint64_t n, Fq2 g0x, Fq2 g0y,
                                                       Fq2 g1x, Fq2 g1y, uint32_t* __restrict__ div,
                                                       uint8_t* __restrict__ flags) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k[8];
    for (int q = 0; q < 8; ++q) k[q] = z[i * 8 + q];
    uint32_t* o = div + i * 32;
    if (!is_canonical<FrParams>(k)) {
        flags[i] = 2;
        for (int q = 0; q < 32; ++q) o[q] = 0;
        return;
    }
    flags[i] = 0;
    const Xyzz<Fq2Ops> g0 = {g0x, g0y, f2_one(), f2_one()};
    const Xyzz<Fq2Ops> zg = xyzz_mul_words<Fq2Ops, 8>(g0, k, 254);
    const Affine<Fq2Ops> d = xyzz_to_affine<Fq2Ops>(xyzz_add<Fq2Ops>({g1x, g1y, f2_one(), f2_one()},
                                                                       xyzz_neg<Fq2Ops>(zg)));
    const Fq* c[4] = {&d.x.c0, &d.x.c1, &d.y.c0, &d.y.c1};
    for (int q = 0; q < 4; ++q) {
        uint32_t w[8];
        pack(d.inf ? fp_zero<FqParams>() : *c[q], w);
        for (int u = 0; u < 8; ++u) o[q * 8 + u] = w[u];
    }
}

struct GlobalLines {
    const uint32_t* p;
    __device__ Line operator()(int i) const {
        const uint32_t* q = p + i * 48;
        return {{ldfq(q), ldfq(q + 8)}, {ldfq(q + 16), ldfq(q + 24)}, {ldfq(q + 32), ldfq(q + 40)}};
    }
};

// Check (w, i): e(C_w - [v_wi] g1_0, g2_0) == e(pi_wi, D_i), as the product
// e(C_w - [v_wi] g1_0, g2_0) e(-pi_wi, D_i) == 1; g2_0's lines prepared.
__global__ void __launch_bounds__(64) k_cache_check(const uint32_t* __restrict__ values, const uint32_t* __restrict__ caches,
                                                    const uint32_t* __restrict__ div, const uint8_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ commitments, uint32_t wires, uint64_t n,
                                                    Fq g1x, Fq g1y, const uint32_t* __restrict__ lines,
                                                    uint8_t* __restrict__ ok) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= uint64_t(wires) * n) return;
    const uint64_t w = t / n, i = t % n;
    const uint32_t* v = values + t * 8;
    const uint32_t* pi = caches + t * 16;
    bool valid = flags[i] == 0 && canonical_fr(v);
    const bool pi0 = all_zero(pi, 16);
    Fq px = fp_zero<FqParams>(), py = fp_zero<FqParams>();
    if (!pi0) {
        valid = valid && canonical_fq(pi) && canonical_fq(pi + 8);
        if (valid) {
            px = ld_canon_fq(pi);
            py = ld_canon_fq(pi + 8);
            valid = affine_on_curve<FqOps>({px, py, false});
        }
    }
    if (!valid) {
        ok[t] = 2;
        return;
    }
    uint32_t k[8];
    for (int q = 0; q < 8; ++q) k[q] = v[q];
    const uint32_t* cw = commitments + w * 16;  // Montgomery affine, all-zero = infinity
    const G1 c = all_zero(cw, 16) ? xyzz_inf<FqOps>()
                                  : G1{ldfq(cw), ldfq(cw + 8), fp_one<FqParams>(), fp_one<FqParams>()};
    const G1 vg = g1_mul_scalar({g1x, g1y, fp_one<FqParams>(), fp_one<FqParams>()}, k, false);
    const Affine<FqOps> d = xyzz_to_affine<FqOps>(xyzz_add<FqOps>(c, xyzz_neg<FqOps>(vg)));
    const uint32_t* dv = div + i * 32;
    const bool dinf = all_zero(dv, 32);
    PairingInput in[1];
    int m = 0;
    if (!pi0 && !dinf) in[m++] = {px, neg(py), {ldfq(dv), ldfq(dv + 8)}, {ldfq(dv + 16), ldfq(dv + 24)}};
    const Fq12 f = (m || !d.inf) ? miller_product_impl(in, m, GlobalLines{lines}, !d.inf, d.x, d.y) : f12_one();
    ok[t] = f12_is_one(final_exponentiation(f)) ? 1 : 0;
}

// KZG opening check of item t (one proof each):
//     e(C_t - [y_t] g1_0, g2_0) == e(W_t, D_t),
// as the product e(C_t - [y_t] g1_0, g2_0) e(-W_t, D_t) == 1, g2_0's lines
// prepared.  Every input is canonical on the wire (C_t, W_t 64 B; y_t 32 B;
// D_t 128 B) and checked here: ok[t] = 2 flags a bad one for the host.
__global__ void __launch_bounds__(64) k_kzg_verify(const uint32_t* __restrict__ commitments,
                                                   const uint32_t* __restrict__ values,
                                                   const uint32_t* __restrict__ openings,
                                                   const uint32_t* __restrict__ divisors, uint64_t count, Fq g1x,
                                                   Fq g1y, const uint32_t* __restrict__ lines,
                                                   uint8_t* __restrict__ ok) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t* cw = commitments + t * 16;
    const uint32_t* v = values + t * 8;
    const uint32_t* pi = openings + t * 16;
    const uint32_t* dv = divisors + t * 32;
    bool valid = canonical_fr(v);
    const bool c0 = all_zero(cw, 16), pi0 = all_zero(pi, 16), dinf = all_zero(dv, 32);
    G1 c = xyzz_inf<FqOps>();
    if (!c0) {
        valid = valid && canonical_fq(cw) && canonical_fq(cw + 8);
        if (valid) {
            const Fq x = ld_canon_fq(cw), y = ld_canon_fq(cw + 8);
            valid = affine_on_curve<FqOps>({x, y, false});
            c = G1{x, y, fp_one<FqParams>(), fp_one<FqParams>()};
        }
    }
    Fq px = fp_zero<FqParams>(), py = fp_zero<FqParams>();
    if (!pi0) {
        valid = valid && canonical_fq(pi) && canonical_fq(pi + 8);
        if (valid) {
            px = ld_canon_fq(pi);
            py = ld_canon_fq(pi + 8);
            valid = affine_on_curve<FqOps>({px, py, false});
        }
    }
    Fq2 dx = f2_zero(), dy = f2_zero();
    if (!dinf) {
        for (int q = 0; q < 4; ++q) valid = valid && canonical_fq(dv + 8 * q);
        if (valid) {
            dx = {ld_canon_fq(dv), ld_canon_fq(dv + 8)};
            dy = {ld_canon_fq(dv + 16), ld_canon_fq(dv + 24)};
            valid = affine_on_curve<Fq2Ops>({dx, dy, false});
        }
    }
    if (!valid) {
        ok[t] = 2;
        return;
    }
    uint32_t k[8];
    for (int q = 0; q < 8; ++q) k[q] = v[q];
    const G1 vg = g1_mul_scalar({g1x, g1y, fp_one<FqParams>(), fp_one<FqParams>()}, k, false);
    const Affine<FqOps> d = xyzz_to_affine<FqOps>(xyzz_add<FqOps>(c, xyzz_neg<FqOps>(vg)));
    PairingInput in[1];
    int m = 0;
    if (!pi0 && !dinf) in[m++] = {px, neg(py), dx, dy};
    const Fq12 f = (m || !d.inf) ? miller_product_impl(in, m, GlobalLines{lines}, !d.inf, d.x, d.y) : f12_one();
    ok[t] = f12_is_one(final_exponentiation(f)) ? 1 : 0;
}

// ---- wave-per-check KZG verification ---------------------------------------

constexpr uint32_t LINE_WORDS = PAIRING_LINES * 48;

// Canonical wire G2 point (32 words) -> affine Montgomery; false when it is
// the identity, not canonical or off the twist.
__device__ __forceinline__ bool load_g2(const uint32_t* q, Fq2& x, Fq2& y) {
    if (all_zero(q, 32)) return false;
    for (int k = 0; k < 4; ++k)
        if (!canonical_fq(q + 8 * k)) return false;
    x = {ld_canon_fq(q), ld_canon_fq(q + 8)};
    y = {ld_canon_fq(q + 16), ld_canon_fq(q + 24)};
    return affine_on_curve<Fq2Ops>({x, y, false});
}

__device__ __forceinline__ void store_line(uint32_t* out, const Line& l) {
    const Fq* c[6] = {&l.l0.c0, &l.l0.c1, &l.l1.c0, &l.l1.c1, &l.l3.c0, &l.l3.c1};
    for (int q = 0; q < 6; ++q) stfq(out + q * 8, *c[q]);
}

// One lane per G2 point (canonical wire words): its prepared lines, in the
// order of prepare_lines, at lines + slots[i] * LINE_WORDS.  Points that do
// not parse are skipped (their checks fail on their own validation).
__global__ void __launch_bounds__(64) k_lines_many(const uint32_t* __restrict__ wire, uint32_t count,
                                                   const uint32_t* __restrict__ slots, uint32_t* __restrict__ lines) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    Fq2 x2, y2;
    if (!load_g2(wire + size_t(i) * 32, x2, y2)) return;
    uint32_t* out = lines + size_t(slots[i]) * LINE_WORDS;
    G2Proj t = {x2, y2, f2_one()};
    int n = 0;
    for (int b = PAIRING_ATE_BITS - 2; b >= 0; --b) {
        store_line(out + 48 * n++, miller_dbl_coeffs(t));
        if (ate_bit(b)) store_line(out + 48 * n++, miller_add_coeffs(t, x2, y2));
    }
    Fq2 q1x, q1y, q2x, q2y;
    twist_frobenius(x2, y2, q1x, q1y);
    twist_frobenius(q1x, q1y, q2x, q2y);
    store_line(out + 48 * n++, miller_add_coeffs(t, q1x, q1y));
    store_line(out + 48 * n++, miller_add_coeffs(t, q2x, f2_neg(q2y)));
}

// [2^i] P for i < 256 as XYZZ words (one lane).
__global__ void k_g1_doublings(Fq x, Fq y, uint32_t* __restrict__ out) {
    if (blockIdx.x || threadIdx.x) return;
    G1 p = {x, y, fp_one<FqParams>(), fp_one<FqParams>()};
    for (int i = 0; i < 256; ++i) {
        stg1(out + 32 * i, p);
        p = xyzz_dbl<FqOps>(p);
    }
}

__device__ __forceinline__ G1 shfl_down_g1(const G1& p, int off) {
    G1 r;
    const Fq* src[4] = {&p.X, &p.Y, &p.ZZ, &p.ZZZ};
    Fq* dst[4] = {&r.X, &r.Y, &r.ZZ, &r.ZZZ};
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < LIMBS; ++i) dst[k]->l[i] = __shfl_down(src[k]->l[i], off, 64);
    return r;
}

// One team (one wave) per item t (the check of k_kzg_verify): lane 0
// validates the inputs; the wave forms [y_t] g1_0 from the doubling table
// (four bits a lane, then a shuffle tree); the pair (C_t - [y_t] g1_0, g2_0) is taken in
// XYZZ form, its lines scaled by Y ZZ, X ZZZ and ZZ ZZZ (the affine line times
// ZZ ZZZ, a factor in Fq that the final exponentiation removes), the pair
// (-W_t, D_t) by -y, x and 1; both lines come prepared (g2_slot and
// div_slots[t] in `lines`).  Then the team Miller product and final
// exponentiation.
template <int WAVES>
__global__ void __launch_bounds__(TEAM_THREADS, WAVES) k_kzg_team(const uint32_t* __restrict__ commitments,
                                                 const uint32_t* __restrict__ values,
                                                 const uint32_t* __restrict__ openings,
                                                 const uint32_t* __restrict__ divisors,
                                                 const uint32_t* __restrict__ g1_table,
                                                 const uint32_t* __restrict__ lines, uint32_t g2_slot,
                                                 const uint32_t* __restrict__ div_slots,
                                                 const uint32_t* __restrict__ programs,
                                                 const uint32_t* __restrict__ prog_off, uint8_t* __restrict__ ok) {
    TeamShared& s = team_lds;
    const uint64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t* cw = commitments + t * 16;
    const uint32_t* v = values + t * 8;
    const uint32_t* pi = openings + t * 16;
    const uint32_t* dv = divisors + t * 32;
    __shared__ G1 c_point;
    if (lane == 0) {
        bool valid = canonical_fr(v);
        const bool c0 = all_zero(cw, 16), pi0 = all_zero(pi, 16), dinf = all_zero(dv, 32);
        G1 c = xyzz_inf<FqOps>();
        if (!c0) {
            valid = valid && canonical_fq(cw) && canonical_fq(cw + 8);
            if (valid) {
                const Fq x = ld_canon_fq(cw), y = ld_canon_fq(cw + 8);
                valid = affine_on_curve<FqOps>({x, y, false});
                c = G1{x, y, fp_one<FqParams>(), fp_one<FqParams>()};
            }
        }
        Fq px = fp_zero<FqParams>(), py = fp_zero<FqParams>();
        if (!pi0) {
            valid = valid && canonical_fq(pi) && canonical_fq(pi + 8);
            if (valid) {
                px = ld_canon_fq(pi);
                py = ld_canon_fq(pi + 8);
                valid = affine_on_curve<FqOps>({px, py, false});
            }
        }
        if (!dinf) {
            Fq2 dx, dy;
            valid = valid && load_g2(dv, dx, dy);
        }
        c_point = c;
        s.scale[1][0] = {neg(py), fp_zero<FqParams>()};
        s.scale[1][1] = {px, fp_zero<FqParams>()};
        s.scale[1][2] = f2_one();
        s.flags[0] = valid;
        s.live[1] = !pi0 && !dinf;
    }
    team_sync();
    if (!s.flags[0]) {
        if (lane == 0) ok[t] = 2;
        return;
    }
    // [y] g1_0: lane L adds the table points of bits 4L .. 4L + 3, then a
    // shuffle tree
    G1 acc = xyzz_inf<FqOps>();
    if (lane < 64) {
        const uint32_t nib = (v[lane >> 3] >> ((lane & 7) * 4)) & 15u;
        for (int b = 0; b < 4; ++b)
            if ((nib >> b) & 1u) acc = xyzz_add<FqOps>(acc, ldg1(g1_table + 32 * (4 * lane + b)));
        for (int off = 32; off > 0; off >>= 1) acc = xyzz_add<FqOps>(acc, shfl_down_g1(acc, off));
    }
    if (lane == 0) {
        const G1 d = xyzz_add<FqOps>(c_point, xyzz_neg<FqOps>(acc));
        const bool live = !xyzz_is_inf<FqOps>(d);
        if (live) {
            s.scale[0][0] = {mul(d.Y, d.ZZ), fp_zero<FqParams>()};
            s.scale[0][1] = {mul(d.X, d.ZZZ), fp_zero<FqParams>()};
            s.scale[0][2] = {mul(d.ZZ, d.ZZZ), fp_zero<FqParams>()};
        }
        s.live[0] = live;
        s.lines[0] = lines + size_t(g2_slot) * LINE_WORDS;
        s.lines[1] = s.live[1] ? lines + size_t(div_slots[t]) * LINE_WORDS : lines;
    }
    team_clear_lines(s);
    team_run(s, programs + prog_off[4], int(prog_off[5]));  // the two-pair program
    if (lane == 0) ok[t] = f12_is_one(w12_to_tower(s.w[TW_OUT])) ? 1 : 0;
}

// One team per check of `count` groups of pairs: group t holds the pairs
// first[t] .. first[t + 1] - 1 (g1 64 B, g2 128 B canonical wire words, G2
// lines prepared at slots[pair]); in product form, pairs after the first get
// their G1 point negated when `negate_rest` (an equation e(L) == prod e(R_j)
// as one product).  Lane 0 validates every pair (ok[t] = 2 on a bad point);
// a pair with an identity contributes one.  The program of the group's pair
// count comes from programs + prog_off[k], prog_len[k].  gt (optional):
// the reduced product, tower order, canonical words.
template <int WAVES>
__global__ void __launch_bounds__(TEAM_THREADS, WAVES) k_pairs_team(const uint32_t* __restrict__ g1,
                                                             const uint32_t* __restrict__ g2,
                                                             const uint32_t* __restrict__ first,
                                                             const uint32_t* __restrict__ order, uint32_t terms,
                                                             int negate_rest, const uint32_t* __restrict__ slots,
                                                             const uint32_t* __restrict__ lines,
                                                             const uint32_t* __restrict__ programs,
                                                             const uint32_t* __restrict__ prog_off,
                                                             uint8_t* __restrict__ ok, uint32_t* __restrict__ gt) {
    TeamShared& s = team_lds;
    const uint64_t t = order ? order[blockIdx.x] : blockIdx.x;
    const int lane = threadIdx.x;
    const uint64_t p0 = first ? first[t] : t * terms;
    const int k = int(first ? first[t + 1] - first[t] : terms);
    if (lane == 0) {
        bool valid = true;
        for (int j = 0; j < k; ++j) {
            PairingInput e;
            bool finite;
            valid = load_pair(g1 + (p0 + j) * 16, g2 + (p0 + j) * 32, e, finite) && valid;
            s.live[j] = finite;
            if (finite) {
                s.scale[j][0] = {negate_rest && j ? neg(e.yp) : e.yp, fp_zero<FqParams>()};
                s.scale[j][1] = {e.xp, fp_zero<FqParams>()};
                s.scale[j][2] = f2_one();
                s.lines[j] = lines + size_t(slots[p0 + j]) * LINE_WORDS;
            }
        }
        s.flags[0] = valid;
    }
    team_clear_lines(s);
    if (!s.flags[0]) {
        if (lane == 0) ok[t] = 2;
        return;
    }
    team_run(s, programs + prog_off[2 * k], int(prog_off[2 * k + 1]));
    if (lane == 0) {
        const Fq12 r = w12_to_tower(s.w[TW_OUT]);
        ok[t] = f12_is_one(r) ? 1 : 0;
        if (gt) {
            uint32_t* o = gt + t * 96;
            const Fq6* c[2] = {&r.c0, &r.c1};
            for (int h = 0; h < 2; ++h) {
                const Fq2* a[3] = {&c[h]->a0, &c[h]->a1, &c[h]->a2};
                for (int i = 0; i < 3; ++i) {
                    st_canon_fq(o + (h * 3 + i) * 16, a[i]->c0);
                    st_canon_fq(o + (h * 3 + i) * 16 + 8, a[i]->c1);
                }
            }
        }
    }
}

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

Affine<FqOps> host_g1(const uint8_t* b, const char* what) {
    uint32_t x[8], y[8];
    std::memcpy(x, b, 32);
    std::memcpy(y, b + 32, 32);
    bool any = false;
    for (int q = 0; q < 8; ++q) any |= (x[q] | y[q]) != 0;
    if (!any) return {fp_zero<FqParams>(), fp_zero<FqParams>(), true};
    require(is_canonical<FqParams>(x) && is_canonical<FqParams>(y), what);
    const Affine<FqOps> a = {to_mont<FqParams>(x), to_mont<FqParams>(y), false};
    require(affine_on_curve<FqOps>(a), what);
    return a;
}

Affine<Fq2Ops> host_g2(const uint8_t* b, const char* what) {
    uint32_t w[4][8];
    bool any = false;
    for (int k = 0; k < 4; ++k) {
        std::memcpy(w[k], b + 32 * k, 32);
        for (int q = 0; q < 8; ++q) any |= w[k][q] != 0;
    }
    require(any, what);
    for (int k = 0; k < 4; ++k) require(is_canonical<FqParams>(w[k]), what);
    const Affine<Fq2Ops> a = {{to_mont<FqParams>(w[0]), to_mont<FqParams>(w[1])},
                              {to_mont<FqParams>(w[2]), to_mont<FqParams>(w[3])}, false};
    require(affine_on_curve<Fq2Ops>(a), what);
    return a;
}

PairingCache& pairing_cache(Owner* o) {
    if (!o->pairing) o->pairing = new PairingCache();
    return *o->pairing;
}

// Slots of the wire G2 points pts[i] (128 bytes each; at most LINE_SLOTS
// distinct ones); points without lines are prepared on `st`.
std::vector<uint32_t> line_slots(Owner* o, const std::vector<const uint8_t*>& pts, hipStream_t st) {
    PairingCache& c = pairing_cache(o);
    c.lines.ensure(size_t(LINE_SLOTS) * LINE_WORDS * 4);
    std::unordered_map<std::string, uint32_t> fresh;  // points of this call without lines
    for (const uint8_t* p : pts) {
        std::string key(reinterpret_cast<const char*>(p), 128);
        if (!c.slot_of.count(key)) fresh.emplace(std::move(key), 0u);
    }
    if (c.slot_of.size() + fresh.size() > LINE_SLOTS) {  // start over: every point of the call is fresh
        c.slot_of.clear();
        fresh.clear();
        for (const uint8_t* p : pts) fresh.emplace(std::string(reinterpret_cast<const char*>(p), 128), 0u);
    }
    require(fresh.size() <= LINE_SLOTS, "too many distinct G2 points for the prepared-line cache");
    if (!fresh.empty()) {
        std::vector<uint32_t> host(fresh.size() * 33);
        uint32_t* words = host.data();
        uint32_t* slots = host.data() + fresh.size() * 32;
        size_t k = 0;
        for (auto& kv : fresh) {
            const uint32_t slot = uint32_t(c.slot_of.size());
            c.slot_of.emplace(kv.first, slot);
            std::memcpy(words + 32 * k, kv.first.data(), 128);
            slots[k++] = slot;
        }
        auto* dev = static_cast<uint32_t*>(c.staged.ensure(host.size() * 4));
        HIP_CHECK(hipMemcpyAsync(dev, host.data(), host.size() * 4, hipMemcpyHostToDevice, st));
        k_lines_many<<<grid_for(k, 64), 64, 0, st>>>(dev, uint32_t(k), dev + 32 * k, static_cast<uint32_t*>(c.lines.ptr));
        HIP_CHECK_LAUNCH();
        // the pageable source must outlive the copy
        HIP_CHECK(hipStreamSynchronize(st));
    }
    std::vector<uint32_t> out(pts.size());
    for (size_t i = 0; i < pts.size(); ++i)
        out[i] = c.slot_of.at(std::string(reinterpret_cast<const char*>(pts[i]), 128));
    return out;
}

// The team programs of checks of 1..PAIRING_MAX_TERMS pairs, uploaded once
// per owner; returns the programs and sets off to the table of (offset,
// length) pairs indexed by 2 * pairs.
const uint32_t* team_programs(Owner* o, const uint32_t*& off, hipStream_t st) {
    PairingCache& c = pairing_cache(o);
    if (!c.program_words) {
        std::vector<uint32_t> table(2 * (PAIRING_MAX_TERMS + 1), 0u), prog;
        for (int k = 1; k <= PAIRING_MAX_TERMS; ++k) {
            const size_t at = prog.size(), len = size_t(team_program(k, nullptr));
            prog.resize(at + len);
            team_program(k, prog.data() + at);
            table[2 * k] = uint32_t(at);
            table[2 * k + 1] = uint32_t(len);
        }
        const size_t words = table.size() + prog.size();
        std::vector<uint32_t> all(prog);
        all.insert(all.end(), table.begin(), table.end());
        auto* dev = static_cast<uint32_t*>(c.program.ensure(words * 4));
        HIP_CHECK(hipMemcpyAsync(dev, all.data(), words * 4, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
        c.program_words = uint32_t(prog.size());
    }
    const uint32_t* base = static_cast<const uint32_t*>(c.program.ptr);
    off = base + c.program_words;
    return base;
}

// The doubling table of the wire G1 point g (canonical, finite).
const uint32_t* g1_doublings(Owner* o, const uint8_t* g, const Affine<FqOps>& a, hipStream_t st) {
    PairingCache& c = pairing_cache(o);
    const std::string key(reinterpret_cast<const char*>(g), 64);
    if (c.g1_key != key || !c.g1_table.ptr) {
        c.g1_key.clear();
        c.g1_table.ensure(256 * 32 * 4);
        k_g1_doublings<<<1, 64, 0, st>>>(a.x, a.y, static_cast<uint32_t*>(c.g1_table.ptr));
        HIP_CHECK_LAUNCH();
        c.g1_key = key;
    }
    return static_cast<const uint32_t*>(c.g1_table.ptr);
}

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

namespace {

// Team kernels are built for one and for two waves per SIMD: two hold the
// registers to 128 per lane at the cost of ~8% more time per check, which
// pays once a call has more checks than the device has SIMDs (MI355X: 1024).
constexpr size_t TEAM_ONE_WAVE_MAX = 1024;

// Up to this many items a call runs one wave per item (k_kzg_team); above,
// one lane per item (k_kzg_verify), whose 64 items per wave make up for its
// serial latency once the device is full.
constexpr size_t KZG_TEAM_MAX = 8192;

int kzg_verify_team(Owner* o, const uint8_t* g1_0, const uint8_t* g2_0, const Affine<FqOps>& g1,
                    const uint8_t* commitments, const uint8_t* values, const uint8_t* openings,
                    const uint8_t* divisors, size_t count, uint8_t* results, float* kernel_ms) {
    hipStream_t st = o->stream;
    const uint32_t* table = g1_doublings(o, g1_0, g1, st);
    const uint32_t* prog_off = nullptr;
    const uint32_t* programs = team_programs(o, prog_off, st);
    const size_t bc = count * 64, bv = count * 32, bw = count * 64, bd = count * 128,
                 bs = (count * 4 + 255) & ~size_t(255), br = (count + 255) & ~size_t(255);
    auto* base = static_cast<uint8_t*>(o->stage_a.ensure(bc + bv + bw + bd + bs + br));
    uint8_t *dc = base, *dv = dc + bc, *dw = dv + bv, *dd = dw + bw, *ds = dd + bd, *dr = ds + bs;
    HIP_CHECK(hipMemcpyAsync(dc, commitments, bc, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(dv, values, bv, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(dw, openings, bw, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(dd, divisors, bd, hipMemcpyHostToDevice, st));
    ScopedEvents<2> ev(kernel_ms != nullptr);
    if (kernel_ms) HIP_CHECK(hipEventRecord(ev[0], st));
    std::vector<uint8_t> zero(128, 0);
    // chunks whose G2 points (g2_0 and the finite divisors) fit the line cache
    for (size_t first = 0; first < count;) {
        std::vector<const uint8_t*> pts = {g2_0};
        std::vector<size_t> idx;  // pts index of each item's divisor (0: none)
        std::unordered_map<std::string, size_t> seen;
        size_t last = first;
        for (; last < count; ++last) {
            const uint8_t* d = divisors + 128 * last;
            if (std::memcmp(d, zero.data(), 128) == 0) {
                idx.push_back(0);
                continue;
            }
            std::string key(reinterpret_cast<const char*>(d), 128);
            auto it = seen.find(key);
            if (it == seen.end()) {
                if (pts.size() == LINE_SLOTS) break;
                it = seen.emplace(std::move(key), pts.size()).first;
                pts.push_back(d);
            }
            idx.push_back(it->second);
        }
        const std::vector<uint32_t> slots = line_slots(o, pts, st);
        std::vector<uint32_t> item_slots(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) item_slots[k] = slots[idx[k]];
        uint32_t* dslots = reinterpret_cast<uint32_t*>(ds) + first;
        HIP_CHECK(hipMemcpyAsync(dslots, item_slots.data(), item_slots.size() * 4, hipMemcpyHostToDevice, st));
        const uint32_t* lines = static_cast<const uint32_t*>(pairing_cache(o).lines.ptr);
        auto* kernel = last - first > TEAM_ONE_WAVE_MAX ? k_kzg_team<2> : k_kzg_team<1>;
        kernel<<<unsigned(last - first), TEAM_THREADS, 0, st>>>(
            reinterpret_cast<const uint32_t*>(dc) + first * 16, reinterpret_cast<const uint32_t*>(dv) + first * 8,
            reinterpret_cast<const uint32_t*>(dw) + first * 16, reinterpret_cast<const uint32_t*>(dd) + first * 32,
            table, lines, slots[0], dslots, programs, prog_off, dr + first);
        HIP_CHECK_LAUNCH();
        // the pageable slot array must outlive its copy
        HIP_CHECK(hipStreamSynchronize(st));
        first = last;
    }
    if (kernel_ms) HIP_CHECK(hipEventRecord(ev[1], st));
    HIP_CHECK(hipMemcpyAsync(results, dr, count, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (kernel_ms) *kernel_ms = ev.ms(0, 1);
    const void* bad = std::memchr(results, 2, count);
    if (bad) {
        const size_t t = static_cast<const uint8_t*>(bad) - results;
        std::memset(results, 0, count);
        throw InvalidArgument(("KZG opening " + std::to_string(t) +
                               ": a point or value is not canonical or not on its curve").c_str());
    }
    return 0;
}

// Up to this many checks (pairing_check, pairing_equations in product mode)
// run one team each, when their distinct G2 points fit the line cache.
constexpr size_t PAIRS_TEAM_MAX = 4096;

// Line-cache slots of the n wire G2 points at g2 (128 bytes each; identity
// points get slot 0, never read), or an empty vector when the distinct
// points do not fit the cache.
std::vector<uint32_t> pair_slots(Owner* o, const uint8_t* g2, size_t n, hipStream_t st) {
    static const uint8_t zero[128] = {0};
    std::unordered_map<std::string, size_t> seen;
    std::vector<const uint8_t*> pts;
    std::vector<size_t> idx(n, SIZE_MAX);
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* q = g2 + 128 * i;
        if (std::memcmp(q, zero, 128) == 0) continue;
        auto it = seen.emplace(std::string(reinterpret_cast<const char*>(q), 128), pts.size()).first;
        if (it->second == pts.size()) pts.push_back(q);
        idx[i] = it->second;
        if (pts.size() > LINE_SLOTS) return {};
    }
    const std::vector<uint32_t> slots = pts.empty() ? std::vector<uint32_t>() : line_slots(o, pts, st);
    std::vector<uint32_t> out(n, 0u);
    for (size_t i = 0; i < n; ++i)
        if (idx[i] != SIZE_MAX) out[i] = slots[idx[i]];
    return out;
}

// The team kernel over count groups already on the device (g1, g2 at dg1,
// dg2; group boundaries at dfirst or `terms` pairs each); results at dres.
void pairs_team_launch(Owner* o, const uint32_t* dg1, const uint32_t* dg2, const uint32_t* dfirst, uint32_t terms,
                       int negate_rest, const std::vector<uint32_t>& slots, size_t count, uint8_t* dres,
                       uint32_t* dgt, hipStream_t st) {
    PairingCache& c = pairing_cache(o);
    const uint32_t* prog_off = nullptr;
    const uint32_t* programs = team_programs(o, prog_off, st);
    auto* dslots = static_cast<uint32_t*>(c.staged.ensure(std::max<size_t>(slots.size(), 1) * 4));
    HIP_CHECK(hipMemcpyAsync(dslots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, st));
    auto* kernel = count > TEAM_ONE_WAVE_MAX ? k_pairs_team<2> : k_pairs_team<1>;
    kernel<<<unsigned(count), TEAM_THREADS, 0, st>>>(dg1, dg2, dfirst, nullptr, terms, negate_rest, dslots,
                                                    static_cast<const uint32_t*>(c.lines.ptr), programs, prog_off,
                                                    dres, dgt);
    HIP_CHECK_LAUNCH();
    // the pageable slot array must outlive its copy
    HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace

void dvgpu::pairing_cache_destroy(PairingCache* c) noexcept { delete c; }

extern "C" int dvgpu_bn254_pairing_check(void* backend, const uint8_t* g1, const uint8_t* g2, size_t terms,
                                         size_t count, uint8_t* results, uint8_t* gt, float* kernel_ms,
                                         char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(terms >= 1 && terms <= PAIRING_MAX_TERMS, "terms must be 1..16");
        require(count <= (size_t(1) << 24), "too many pairing checks");
        if (!count) return 0;
        require(g1 && g2 && results, "null pairing buffers");
        const size_t n = terms * count;
        hipStream_t st = o->stream;
        // one device allocation: g1 | g2 | results | gt
        const size_t b1 = n * 64, b2 = n * 128, br = (count + 255) & ~size_t(255), bg = gt ? count * 384 : 0;
        auto* base = static_cast<uint8_t*>(o->stage_a.ensure(b1 + b2 + br + bg));
        HIP_CHECK(hipMemcpyAsync(base, g1, b1, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(base + b1, g2, b2, hipMemcpyHostToDevice, st));
        ScopedEvents<2> ev(kernel_ms != nullptr);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[0], st));
        const std::vector<uint32_t> slots = count <= PAIRS_TEAM_MAX ? pair_slots(o, g2, n, st) : std::vector<uint32_t>();
        if (!slots.empty())
            pairs_team_launch(o, reinterpret_cast<const uint32_t*>(base), reinterpret_cast<const uint32_t*>(base + b1),
                              nullptr, uint32_t(terms), 0, slots, count, base + b1 + b2,
                              gt ? reinterpret_cast<uint32_t*>(base + b1 + b2 + br) : nullptr, st);
        else
            k_pairing_check<<<grid_for(count, 64), 64, 0, st>>>(
                reinterpret_cast<const uint32_t*>(base), reinterpret_cast<const uint32_t*>(base + b1), uint32_t(terms),
                count, base + b1 + b2, gt ? reinterpret_cast<uint32_t*>(base + b1 + b2 + br) : nullptr);
        HIP_CHECK_LAUNCH();
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[1], st));
        HIP_CHECK(hipMemcpyAsync(results, base + b1 + b2, count, hipMemcpyDeviceToHost, st));
        if (gt) HIP_CHECK(hipMemcpyAsync(gt, base + b1 + b2 + br, bg, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (kernel_ms) *kernel_ms = ev.ms(0, 1);
        const void* bad = std::memchr(results, 2, count);
        if (bad) {
            const size_t i = static_cast<const uint8_t*>(bad) - results;
            std::memset(results, 0, count);
            if (gt) std::memset(gt, 0, count * 384);
            throw InvalidArgument(("pairing check " + std::to_string(i) +
                                   ": a point is not canonical or not on its curve").c_str());
        }
        return 0;
    });
}

extern "C" int dvgpu_bn254_pairing_equations(void* backend, const uint8_t* g1, const uint8_t* g2,
                                             const uint32_t* offsets, size_t count, int mode, uint8_t* results,
                                             float* kernel_ms, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(mode == DVGPU_PAIRING_PRODUCT || mode == DVGPU_PAIRING_REFERENCE, "unknown verification mode");
        require(count <= (size_t(1) << 24), "too many pairing equations");
        if (!count) return 0;
        require(offsets && results, "null pairing-equation buffers");
        require(offsets[0] == 0, "equation offsets must start at zero");
        for (size_t i = 0; i < count; ++i) {
            require(offsets[i + 1] > offsets[i], "every equation needs its left pair");
            require(offsets[i + 1] - offsets[i] <= uint32_t(PAIRING_MAX_TERMS), "an equation has more than 16 pairs");
        }
        const size_t n = offsets[count];
        require(n <= (size_t(1) << 24), "too many pairs");
        require(g1 && g2, "null pairing buffers");
        // equations by pair count (stable counting sort), launched in that order
        std::vector<uint32_t> order(count);
        {
            size_t start[PAIRING_MAX_TERMS + 2] = {0};
            for (size_t i = 0; i < count; ++i) ++start[offsets[i + 1] - offsets[i] + 1];
            for (int k = 1; k <= PAIRING_MAX_TERMS + 1; ++k) start[k] += start[k - 1];
            for (size_t i = 0; i < count; ++i) order[start[offsets[i + 1] - offsets[i]]++] = uint32_t(i);
        }
        hipStream_t st = o->stream;
        // one device allocation: g1 | g2 | offsets | order | results
        const size_t b1 = n * 64, b2 = n * 128, bo = ((count + 1) * 4 + 255) & ~size_t(255),
                     bq = (count * 4 + 255) & ~size_t(255);
        auto* base = static_cast<uint8_t*>(o->stage_a.ensure(b1 + b2 + bo + bq + count));
        HIP_CHECK(hipMemcpyAsync(base, g1, b1, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(base + b1, g2, b2, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(base + b1 + b2, offsets, (count + 1) * 4, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(base + b1 + b2 + bo, order.data(), count * 4, hipMemcpyHostToDevice, st));
        ScopedEvents<2> ev(kernel_ms != nullptr);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[0], st));
        const std::vector<uint32_t> slots = mode == DVGPU_PAIRING_PRODUCT && count <= PAIRS_TEAM_MAX
                                                ? pair_slots(o, g2, n, st) : std::vector<uint32_t>();
        if (!slots.empty())  // product form: the right-hand G1 points negated
            pairs_team_launch(o, reinterpret_cast<const uint32_t*>(base), reinterpret_cast<const uint32_t*>(base + b1),
                              reinterpret_cast<const uint32_t*>(base + b1 + b2), 0u, 1, slots, count,
                              base + b1 + b2 + bo + bq, nullptr, st);
        else
            k_pairing_equations<<<grid_for(count, 64), 64, 0, st>>>(
                reinterpret_cast<const uint32_t*>(base), reinterpret_cast<const uint32_t*>(base + b1),
                reinterpret_cast<const uint32_t*>(base + b1 + b2), reinterpret_cast<const uint32_t*>(base + b1 + b2 + bo),
                count, mode == DVGPU_PAIRING_REFERENCE, base + b1 + b2 + bo + bq);
        HIP_CHECK_LAUNCH();
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[1], st));
        HIP_CHECK(hipMemcpyAsync(results, base + b1 + b2 + bo + bq, count, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (kernel_ms) *kernel_ms = ev.ms(0, 1);
        const void* bad = std::memchr(results, 2, count);
        if (bad) {
            const size_t i = static_cast<const uint8_t*>(bad) - results;
            std::memset(results, 0, count);
            throw InvalidArgument(("pairing equation " + std::to_string(i) +
                                   ": a point is not canonical or not on its curve").c_str());
        }
        return 0;
    });
}

extern "C" int dvgpu_bn254_opening_cache_check(void* backend, const uint8_t* commitments, size_t wires,
                                               const uint8_t* g1_0, const uint8_t* g2_0, const uint8_t* g2_1,
                                               const uint8_t* points, size_t n, const uint8_t* values,
                                               const uint8_t* caches, uint8_t* results, float* kernel_ms,
                                               char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(wires >= 1 && wires <= 16, "wires must be 1..16");
        require(n <= (size_t(1) << 22), "too many cache points");
        if (!n) return 0;
        require(commitments && g1_0 && g2_0 && g2_1 && points && values && caches && results,
                "null cache-check buffers");
        const Affine<FqOps> g1 = host_g1(g1_0, "g1 power is not a canonical curve point");
        require(!g1.inf, "g1 power is the identity");
        const Affine<Fq2Ops> h0 = host_g2(g2_0, "g2 power 0 is not a canonical twist point");
        const Affine<Fq2Ops> h1 = host_g2(g2_1, "g2 power 1 is not a canonical twist point");
        std::vector<uint32_t> cm(wires * 16, 0u);
        for (size_t w = 0; w < wires; ++w) {
            const Affine<FqOps> c = host_g1(commitments + 64 * w, "commitment is not a canonical curve point");
            if (c.inf) continue;
            uint32_t a[8], b[8];
            pack(c.x, a);
            pack(c.y, b);
            std::memcpy(cm.data() + 16 * w, a, 32);
            std::memcpy(cm.data() + 16 * w + 8, b, 32);
        }
        hipStream_t st = o->stream;
        const size_t total = wires * n;
        const size_t bz = n * 32, bv = total * 32, bc = total * 64, bd = n * 128, bf = (n + 255) & ~size_t(255),
                     bm = wires * 64, br = (total + 255) & ~size_t(255);
        auto* base = static_cast<uint8_t*>(o->stage_a.ensure(bz + bv + bc + bd + bf + bm + br));
        uint8_t *dz = base, *dv = dz + bz, *dc = dv + bv, *dd = dc + bc, *df = dd + bd, *dm = df + bf, *dr = dm + bm;
        HIP_CHECK(hipMemcpyAsync(dz, points, bz, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(dv, values, bv, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(dc, caches, bc, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(dm, cm.data(), bm, hipMemcpyHostToDevice, st));
        ScopedEvents<2> ev(kernel_ms != nullptr);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[0], st));
        k_cache_divisors<<<grid_for(n, 64), 64, 0, st>>>(reinterpret_cast<const uint32_t*>(dz), n, h0.x, h0.y, h1.x,
                                                          h1.y, reinterpret_cast<uint32_t*>(dd), df);
        HIP_CHECK_LAUNCH();
        // g2_0's lines from the owner's line cache (prepared on first use)
        const uint32_t g2_slot = line_slots(o, std::vector<const uint8_t*>{g2_0}, st)[0];
        const uint32_t* lines = static_cast<const uint32_t*>(pairing_cache(o).lines.ptr) + size_t(g2_slot) * LINE_WORDS;
        k_cache_check<<<grid_for(total, 64), 64, 0, st>>>(
            reinterpret_cast<const uint32_t*>(dv), reinterpret_cast<const uint32_t*>(dc),
            reinterpret_cast<const uint32_t*>(dd), df, reinterpret_cast<const uint32_t*>(dm), uint32_t(wires), n,
            g1.x, g1.y, lines, dr);
        HIP_CHECK_LAUNCH();
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[1], st));
        HIP_CHECK(hipMemcpyAsync(results, dr, total, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (kernel_ms) *kernel_ms = ev.ms(0, 1);
        const void* bad = std::memchr(results, 2, total);
        if (bad) {
            const size_t t = static_cast<const uint8_t*>(bad) - results;
            std::memset(results, 0, total);
            throw InvalidArgument(("cache check wire " + std::to_string(t / n) + " index " + std::to_string(t % n) +
                                   ": a point, value or cache point is not canonical or not on its curve").c_str());
        }
        return 0;
    });
}

extern "C" int dvgpu_bn254_kzg_verify_batch(void* backend, const uint8_t* g1_0, const uint8_t* g2_0,
                                            const uint8_t* commitments, const uint8_t* values,
                                            const uint8_t* openings, const uint8_t* divisors, size_t count,
                                            uint8_t* results, float* kernel_ms, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(count <= (size_t(1) << 24), "too many KZG openings");
        if (!count) return 0;
        require(g1_0 && g2_0 && commitments && values && openings && divisors && results,
                "null KZG verification buffers");
        const Affine<FqOps> g1 = host_g1(g1_0, "g1 generator is not a canonical curve point");
        require(!g1.inf, "g1 generator is the identity");
        const Affine<Fq2Ops> h0 = host_g2(g2_0, "g2 generator is not a canonical twist point");
        require(!h0.inf, "g2 generator is the identity");
        hipStream_t st = o->stream;
        if (count <= KZG_TEAM_MAX) return kzg_verify_team(o, g1_0, g2_0, g1, commitments, values, openings, divisors,
                                                          count, results, kernel_ms);
        const size_t bc = count * 64, bv = count * 32, bw = count * 64, bd = count * 128,
                     br = (count + 255) & ~size_t(255);
        auto* base = static_cast<uint8_t*>(o->stage_a.ensure(bc + bv + bw + bd + br));
        uint8_t *dc = base, *dv = dc + bc, *dw = dv + bv, *dd = dw + bw, *dr = dd + bd;
        HIP_CHECK(hipMemcpyAsync(dc, commitments, bc, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(dv, values, bv, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(dw, openings, bw, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(dd, divisors, bd, hipMemcpyHostToDevice, st));
        ScopedEvents<2> ev(kernel_ms != nullptr);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[0], st));
        // g2_0's lines from the owner's line cache (prepared on first use)
        const uint32_t g2_slot = line_slots(o, std::vector<const uint8_t*>{g2_0}, st)[0];
        const uint32_t* lines = static_cast<const uint32_t*>(pairing_cache(o).lines.ptr) + size_t(g2_slot) * LINE_WORDS;
        k_kzg_verify<<<grid_for(count, 64), 64, 0, st>>>(
            reinterpret_cast<const uint32_t*>(dc), reinterpret_cast<const uint32_t*>(dv),
            reinterpret_cast<const uint32_t*>(dw), reinterpret_cast<const uint32_t*>(dd), count, g1.x, g1.y,
            lines, dr);
        HIP_CHECK_LAUNCH();
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[1], st));
        HIP_CHECK(hipMemcpyAsync(results, dr, count, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (kernel_ms) *kernel_ms = ev.ms(0, 1);
        const void* bad = std::memchr(results, 2, count);
        if (bad) {
            const size_t t = static_cast<const uint8_t*>(bad) - results;
            std::memset(results, 0, count);
            throw InvalidArgument(("KZG opening " + std::to_string(t) +
                                   ": a point or value is not canonical or not on its curve").c_str());
        }
        return 0;
    });
}
