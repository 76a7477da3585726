/*
 * dynaverse_poly.h -- Fr polynomial arithmetic with batched device products.
 *
 * Replaces the host polynomial tools of crates/protocol/src/domain/fast.rs:
 * SubproductTree (new / vanishing / inverse_derivatives / multipoint_eval /
 * interpolate / cauchy_sums, fast.rs:9-139), newton_divide_exact
 * (fast.rs:150-201, 236-240) and fft_multiply_with_backend (fast.rs:
 * 213-234).  The reference issues each product as three separate
 * backend.ntt calls; here all products of one tree level (or one Newton
 * step of every node of a level) run as one fused NTT batch on the device.
 * Results are the same polynomials and values.
 *
 * Fr elements are 32 bytes little-endian canonical; polynomials are
 * coefficient arrays, lowest degree first.  Outputs are trimmed as the
 * reference trims (trailing zeros dropped, at least one coefficient).  Every
 * call takes the 2^20-th root of unity of the radix-2 domains (as for
 * dvgpu_bn254_ntt); products are limited to 2^20 coefficients.  Status codes
 * as in dynaverse_gpu.h; the reference's errors (duplicate points, wrong
 * value counts, division by zero, nonzero remainder) are status -1.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Work of one call: device rounds of batched products, all products, and
 * the products that went to the device (a round of small products -- at most
 * 256 coefficient pairs each, 4096 in total -- is formed on the host). */
typedef struct DvgpuPolyStats {
    uint64_t rounds;
    uint64_t products;
    uint64_t device_products;
} DvgpuPolyStats;

/* fft_multiply_with_backend: left * right (an empty operand gives [0]). */
int dvgpu_poly_mul(void* backend, const uint8_t* root, const uint8_t* left, size_t left_len,
                   const uint8_t* right, size_t right_len, uint8_t* out, size_t out_capacity, size_t* out_len,
                   char* error, size_t error_capacity);

/* newton_divide_exact: numerator / denominator, -1 unless exact. */
int dvgpu_poly_divide_exact(void* backend, const uint8_t* root, const uint8_t* numerator, size_t num_len,
                            const uint8_t* denominator, size_t den_len, uint8_t* out, size_t out_capacity,
                            size_t* out_len, char* error, size_t error_capacity);

/* SubproductTree::new over count distinct points (input order kept; -1 for a
 * duplicate).  The tree belongs to `backend` and is freed with it. */
int dvgpu_poly_tree_create(void* backend, const uint8_t* root, const uint8_t* points, size_t count, void** tree,
                           DvgpuPolyStats* stats, char* error, size_t error_capacity);
int dvgpu_poly_tree_destroy(void* backend, void* tree);

/* prod (X - x_i) and 1 / A'(x_i) in input order. */
int dvgpu_poly_tree_vanishing(void* backend, void* tree, uint8_t* out, size_t out_capacity, size_t* out_len,
                              char* error, size_t error_capacity);
int dvgpu_poly_tree_inverse_derivatives(void* backend, void* tree, uint8_t* out, size_t out_capacity,
                                        size_t* out_len, char* error, size_t error_capacity);

/* multipoint_eval: values (count x 32 B) of the polynomial at the points. */
int dvgpu_poly_tree_eval(void* backend, const uint8_t* root, void* tree, const uint8_t* poly, size_t len,
                         uint8_t* values, DvgpuPolyStats* stats, char* error, size_t error_capacity);

/* interpolate: the polynomial of degree < count through (x_i, values_i). */
int dvgpu_poly_tree_interpolate(void* backend, const uint8_t* root, void* tree, const uint8_t* values,
                                size_t count, uint8_t* out, size_t out_capacity, size_t* out_len,
                                DvgpuPolyStats* stats, char* error, size_t error_capacity);

/* cauchy_sums: sum_{j != i} weights_j / (x_i - x_j) for every i. */
int dvgpu_poly_tree_cauchy_sums(void* backend, const uint8_t* root, void* tree, const uint8_t* weights,
                                size_t count, uint8_t* sums, DvgpuPolyStats* stats, char* error,
                                size_t error_capacity);

#ifdef __cplusplus
}
#endif
