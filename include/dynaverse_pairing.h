/*
 * dynaverse_pairing.h -- batched BN254 pairing-product checks on the device.
 *
 * Replaces the host pairing checks of the opening-cache validation
 * (crates/protocol/src/dynaverse/rebuilding/machine.rs:255-390): every cache
 * point pi_i of a commitment C is accepted when
 *     e(C - [v_i] g1, g2) == e(pi_i, [tau]g2 - [z_i]g2),
 * which the reference's PairingChecks::equal evaluates in Product mode as one
 * two-term Miller product and one final exponentiation per point
 * (machine.rs:333, validate_initial_caches machine.rs:357-382).  Here each
 * check is a group of `terms` pairs (P, Q) and the answer is whether
 * prod e(P_j, Q_j) is one in GT; the caller negates one side (e(A, B) ==
 * e(C, D) is the group (A, B), (-C, D)).  The pairing is the optimal ate
 * pairing with the final exponent (q^12 - 1)/r; accept/reject equals that of
 * any non-degenerate bilinear pairing, arkworks' included.
 *
 * Encodings as in dynaverse_gpu.h: G1 64 B x || y, G2 128 B x.c0 || x.c1 ||
 * y.c0 || y.c1, little-endian canonical, all-zero = identity (the pair then
 * contributes one).  Points must be canonical and lie on E (G1) and on the
 * twist E' (G2), checked on the device; otherwise the call fails with -1 and
 * zeroes results and gt.  G2 subgroup membership is NOT checked here: G2
 * points are taken to be in the r-torsion, as for dvgpu_bn254_g2_msm.  The
 * arkworks typed points these checks replace are subgroup-checked when they
 * are deserialised, so a caller holding untrusted G2 wire bytes runs
 * dvgpu_bn254_g2_validation_inspect (dynaverse_gpu.h) on them first and
 * passes only the points it accepts.  Status codes as in dynaverse_gpu.h.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* count groups of `terms` pairs: g1 count x terms x 64 B, g2 count x terms x
 * 128 B (pair j of group i at index i*terms + j).  results: count bytes, 1
 * when the group's product is one.  gt (optional): count x 384 B reduced
 * products, Fq12 = Fq6[w]/(w^2 - v), Fq6 = Fq2[v]/(v^3 - (9 + u)), written
 * c0.a0, c0.a1, c0.a2, c1.a0, c1.a1, c1.a2 (Fq2 each: real, imaginary),
 * canonical LE32 Fq.  kernel_ms (optional): event-timed device duration.
 * terms in 1..16; count up to 2^24.  No counterpart in the reference ABI. */
int dvgpu_bn254_pairing_check(void* backend, const uint8_t* g1, const uint8_t* g2, size_t terms, size_t count,
                              uint8_t* results, uint8_t* gt, float* kernel_ms, char* error,
                              size_t error_capacity);

/* Verification modes of PairingChecks (crates/protocol/src/types.rs:193-196). */
#define DVGPU_PAIRING_PRODUCT 0
#define DVGPU_PAIRING_REFERENCE 1

/* A batch of independent pairing equations, each as PairingChecks::equal
 * states it (crates/protocol/src/verification.rs:37-53):
 *     e(L.P, L.Q) == prod_j e(R_j.P, R_j.Q),
 * the pairs of equation i at indices offsets[i] .. offsets[i+1]-1 of g1
 * (64 B each) and g2 (128 B each), the left pair first.  offsets holds
 * count + 1 entries, offsets[0] = 0, 1..16 pairs per equation, at most 2^24
 * pairs.  mode DVGPU_PAIRING_PRODUCT evaluates an equation as product_equal
 * does (verification.rs:79-103): one Miller product over the left pair and
 * the right pairs with negated G1 points, one final exponentiation, compared
 * with one.  DVGPU_PAIRING_REFERENCE evaluates it as the Reference branch of
 * equal does (verification.rs:46-53): every pair reduced on its own and the
 * right-hand values multiplied in GT, compared with the left value.
 * equal_prepared (verification.rs:55-77) is the two-pair Product equation
 * (left, generator), (right, divisor).  Equations are never merged: results
 * holds one byte per equation, 1 when it holds.  Replaces the arkworks
 * pairings of every PairingChecks caller (dynaverse.rs:1003-1113,
 * dynamo.rs:415-430, consistency.rs:95-106, rebuilding/machine.rs:309-390);
 * the callers' short-circuit order is the caller's.  Point checks and
 * identities as for dvgpu_bn254_pairing_check; a bad point fails the call
 * with -1 and zeroes results.  kernel_ms (optional): event-timed device
 * duration.  No counterpart in the reference ABI. */
int dvgpu_bn254_pairing_equations(void* backend, const uint8_t* g1, const uint8_t* g2, const uint32_t* offsets,
                                  size_t count, int mode, uint8_t* results, float* kernel_ms, char* error,
                                  size_t error_capacity);

/* The whole opening-cache validation of validate_initial_caches
 * (machine.rs:357-382) and validate_cache_block_prepared (machine.rs:
 * 276-336, Product mode) on the device: for every wire w < wires and point
 * i < n, result[w*n + i] = 1 when
 *     e(C_w - [v_wi] g1_0, g2_0) == e(pi_wi, g2_1 - [z_i] g2_0),
 * with commitments C_w (wires x 64 B, the anchor's s4, s5, s6),
 * g1_0 = g1_powers[0], g2_0 = g2_powers[0], g2_1 = g2_powers[1] (the
 * rebuild SRS powers), points z_i (n x 32 B Fr: the domain's small points),
 * values v_wi (wires x n x 32 B: the witness wires 3..5) and caches pi_wi
 * (wires x n x 64 B: q4, q5, q6).  The divisor of point i is formed once and
 * shared by every wire, as machine.rs:362-368 does.  The reference stops at
 * the first failing (wire, i) in the order of its 4-point blocks; callers
 * scan the results in that order for the same error.  Inputs are checked
 * (canonical, on their curves); a bad one fails the call with -1.
 * wires in 1..16, n up to 2^22.  No counterpart in the reference ABI. */
int dvgpu_bn254_opening_cache_check(void* backend, const uint8_t* commitments, size_t wires, const uint8_t* g1_0,
                                    const uint8_t* g2_0, const uint8_t* g2_1, const uint8_t* points, size_t n,
                                    const uint8_t* values, const uint8_t* caches, uint8_t* results,
                                    float* kernel_ms, char* error, size_t error_capacity);

/* KZG single-point opening checks, one proof per item: item t is accepted
 * (results[t] = 1) when
 *     e(C_t - [y_t] g1_0, g2_0) == e(W_t, D_t),
 * with the commitment C_t and opening W_t (count x 64 B G1), the value
 * y_t = f(z_t) (count x 32 B Fr) and the divisor D_t = [tau]g2 - [z_t]g2
 * (count x 128 B G2, the verifier's frozen G2 term of its session), g1_0 and
 * g2_0 the generators.  This is the pairing verification of a KZG update
 * proof (crates/protocol/src/dynaverse.rs:447-451, the cache equation of
 * rebuilding/machine.rs:321-335) for many proofs of many SRSs at once,
 * evaluated as one two-term Miller product (g2_0's lines prepared once) and
 * one final exponentiation per item.  All inputs are checked on the device
 * (canonical, on E / E'); a bad one fails the call with -1 and zeroes
 * results.  D_t must be in the r-torsion, as for every G2 input here.
 * count up to 2^24.  No counterpart in the reference ABI. */
int dvgpu_bn254_kzg_verify_batch(void* backend, const uint8_t* g1_0, const uint8_t* g2_0,
                                 const uint8_t* commitments, const uint8_t* values, const uint8_t* openings,
                                 const uint8_t* divisors, size_t count, uint8_t* results, float* kernel_ms,
                                 char* error, size_t error_capacity);

#ifdef __cplusplus
}
#endif
