// Subproduct tree resident in HBM (the device form of fast.rs:9-139's
// SubproductTree for larger point sets).  Every level, the reciprocals of
// the remainder tree and the transformed operands stay on the device between
// calls; a call uploads its inputs once, runs its rounds as kernels on the
// owner's stream and downloads its outputs once.  Inputs and outputs are
// canonical little-endian Fr words (8 per element).
#pragma once
#include <cstddef>
#include <cstdint>

namespace dvgpu {

struct Owner;
struct DevTree;

struct DevTreeStats {
    uint64_t rounds = 0;    // batched product passes
    uint64_t products = 0;  // polynomial products in them
};

// Point sets of at least this size get a device tree (below it the host
// rounds of poly.hip are cheaper than the fixed cost of the device rounds).
constexpr size_t DEV_TREE_MIN_POINTS = 64;
constexpr size_t DEV_TREE_MAX_POINTS = size_t(1) << 19;

// k distinct points (the caller checks distinctness).  The owner's NTT
// domain must be prepared; the tree keeps using that root.  Throws
// InvalidArgument("repeated interpolation root") if some A'(x_i) is zero.
DevTree* dev_tree_build(Owner* o, const uint32_t* points, size_t k, DevTreeStats& st);
void dev_tree_destroy(DevTree* t) noexcept;
size_t dev_tree_points(const DevTree* t);

// prod (X - x_i): k + 1 coefficients.
void dev_tree_vanishing(DevTree* t, uint32_t* out);
// 1 / A'(x_i), k values.
void dev_tree_inverse_derivatives(DevTree* t, uint32_t* out);
// Values at the k points of `count` (1 or 2) polynomials of at most k
// coefficients each, all in one remainder tree.
void dev_tree_eval(DevTree* t, const uint32_t* const* polys, const size_t* lens, size_t count,
                   uint32_t* const* values, DevTreeStats& st);
// The polynomial of degree < k through (x_i, v_i): k coefficients (untrimmed).
void dev_tree_interpolate(DevTree* t, const uint32_t* values, uint32_t* coeffs, DevTreeStats& st);
// sum_{j != i} w_j / (x_i - x_j) for every i (fast.rs:118-139), k values.
void dev_tree_cauchy(DevTree* t, const uint32_t* weights, uint32_t* sums, DevTreeStats& st);

}  // namespace dvgpu
