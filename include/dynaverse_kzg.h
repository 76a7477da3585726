/*
 * dynaverse_kzg.h -- KZG opening-cache precompute on the device.
 *
 * Replaces the host computation of KzgParams::precompute_single_openings
 * (crates/protocol/src/kzg.rs:217-275) and its retained SRS transform
 * (`prepared_opening_basis`, kzg.rs:277-300): all single-point openings
 * [(f(tau) - f(z)) / (tau - z)] G1 of a polynomial f for every z of a
 * radix-2 domain (or of a coset of it), as one Fr NTT and three G1
 * transforms.  Called from DynaverseCircuit::precompute_anchor_openings
 * (crates/protocol/src/dynaverse.rs:303-314) for the three opening caches
 * q4, q5, q6, which share one degree and can go in one call.
 *
 * Encodings as in dynaverse_gpu.h: Fr 32 bytes little-endian canonical, G1
 * affine 64 bytes x || y little-endian canonical, all-zero = identity.
 * Status codes as in dynaverse_gpu.h (-1 invalid argument, -2/-3 device
 * failure which retires the backend, -4 retired backend).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Event-timed phases of one precompute call (milliseconds). */
typedef struct DvgpuKzgTimings {
    float scalar_and_basis_ms;   /* upload, scalar NTT, basis transform when rebuilt */
    float convolution_ms;        /* pointwise products and inverse G1 transform */
    float domain_transform_ms;   /* domain G1 transform and affine normalisation */
} DvgpuKzgTimings;

/* An SRS on `backend`: count G1 powers [tau^i] G1 (KzgParams::g1_powers,
 * kzg.rs:107-127) copied to device memory, plus the retained convolution
 * basis of the last degree used (kzg.rs:277-300: one entry, keyed by the
 * degree, rebuilt when the degree changes and never changed by a failed
 * call).  root_2_20: the 2^20-th root of unity arkworks' radix-2 domains use
 * (as for dvgpu_bn254_ntt).  Points must be canonical and on the curve.
 * Calls on one SRS are serialised by a lock held through the whole call
 * (kzg.rs:281-299 holds its lock through construction). */
int dvgpu_kzg_srs_create(void* backend, const uint8_t* g1_powers, size_t count, const uint8_t* root_2_20,
                         void** srs, char* error, size_t error_capacity);
int dvgpu_kzg_srs_destroy(void* backend, void* srs);

/* Degree of the retained basis (0 = none) and how many times a basis has
 * been built: the observable state of kzg.rs:277-300's cache. */
int dvgpu_kzg_srs_basis(void* backend, void* srs, uint64_t* degree, uint64_t* builds);

/* precompute_single_openings (kzg.rs:217-275) for poly_count polynomials of
 * coeff_count coefficients each (row-major, 32 B per coefficient) over the
 * domain of domain_size points (a power of two, at most 2^20), or its coset
 * by coset_offset (32 B, nonzero; NULL for the subgroup itself:
 * Radix2EvaluationDomain::get_coset).  openings: poly_count x domain_size x
 * 64 B in domain order.  Errors as kzg.rs:222-228: coeff_count above the
 * domain size, or above the SRS size; a polynomial of degree 0 (at most one
 * coefficient) gives identities without touching the basis.  The
 * convolution length next_pow2(2 deg - 1) is limited to 2^20 points, and
 * poly_count times the convolution (and the domain) length to 2^24.
 * Calls on one SRS may come from several threads (they run one at a time).
 * If one of them fails on the device and retires the backend, retirement
 * waits for the calls already in progress, and calls still waiting for the
 * SRS return -4.  Destroying an SRS waits for calls in progress on it;
 * creating SRS objects must not overlap other calls on the same backend.
 * timings may be NULL. */
int dvgpu_kzg_precompute_single_openings(void* backend, void* srs, const uint8_t* coefficients,
                                         size_t coeff_count, size_t poly_count, size_t domain_size,
                                         const uint8_t* coset_offset, uint8_t* openings,
                                         DvgpuKzgTimings* timings, char* error, size_t error_capacity);

/* GLV split of count canonical scalars k = k1 + k2 lambda (mod r), lambda the
 * eigenvalue of (x, y) -> (beta x, y) on G1, as the G1 transforms use it.
 * Host computation, no device use; out: 40 B per scalar (|k1| 16 B, |k2|
 * 16 B, sign of k1 and of k2 as 32-bit words).  Returns -1 for a
 * non-canonical scalar.  No counterpart in the reference (test access). */
int dvgpu_ext_glv_split(const uint8_t* scalars, size_t count, uint8_t* out);

/* Joint sparse form of (|k1|, |k2|) of the same split, as the transforms use
 * it when a wave shares the twiddle: out 64 B per scalar, four 128-bit masks
 * (digit of |k1| nonzero, negative; digit of |k2| nonzero, negative), bit l
 * for position l.  Host computation; -1 for a non-canonical scalar.  No
 * counterpart in the reference (test access). */
int dvgpu_ext_glv_jsf(const uint8_t* scalars, size_t count, uint8_t* out);

#ifdef __cplusplus
}
#endif
