// Subproduct tree resident in HBM (declared in poly_tree.hpp).
//
// Shape.  K = 2^L >= k; node (l, i) is the product of X - x_j over the
// leaves j in [i 2^l, min((i + 1) 2^l, k)), a monic polynomial of degree
// d(l, i) <= 2^l (0 for a node past the last point: the polynomial 1).
// Every output (vanishing polynomial, values, interpolant, inverse
// derivatives, Cauchy sums) is fixed by the points and the inputs alone, so
// this shape gives the reference's results (fast.rs:9-139) although its
// tree carries unpaired nodes upward instead.
//
// HBM layout, elements of 8 canonical words (32 B):
//   lev   (L + 1) x K: level l, node i at l K + i 2^l, its d(l, i) non-leading
//         coefficients (the leading 1 is implicit), zeros to 2^l;
//   rh    L x 2K: level l, node i at l 2K + i 2^(l+1): rev(node)^-1 mod
//         x^(2^l), zero-padded to 2^(l+1) and transformed (coefficients for
//         rows of at most 2^SMALL_LOG, where products run by schoolbook);
//   nh    L x 2K: the same rows holding the node's non-leading coefficients;
//   deriv, ideriv: A'(x_i) and 1 / A'(x_i) (A the root node).
// Products of two polynomials of a level run as cyclic convolutions of
// length 2^(l+1), which hold them without wrap-around except where noted.
//
// Rounds (each a batched product pass over every node of a level):
//   build      parent = x^(a+b) + x^a B + x^b A + A B for children x^a + A
//              and x^b + B: one product per pair;
//   Newton     inv_2m = inv_m - x^m (inv_m h mod x^m), h the coefficients
//              m..2m-1 of rev inv_m, taken from a cyclic product of length 2m
//              (its wrap-around only reaches coefficients below m-1); every
//              level's reciprocals in lockstep;
//   remainder  r_child = r_parent mod child by q = rev(rev(r_parent) rcp mod
//              x^qlen) and r_parent - q C (only the low d coefficients: q x^d
//              does not reach them);
//   interpolation  parent = P_A x^b + P_B x^a + P_A B + P_B A, the two
//              products summed in the transformed domain.
#include <algorithm>
#include <cstring>
#include <memory>

#include "g1mem.hpp"
#include "owner.hpp"
#include "poly_tree.hpp"

namespace dvgpu {

namespace {

constexpr int SMALL_LOG = 5;  // cyclic products of length <= 32 by schoolbook
constexpr unsigned BLOCK = 256;

__device__ __forceinline__ Fr ldw(const uint32_t* p) {
    const uint4 a = reinterpret_cast<const uint4*>(p)[0], b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return unpack<FrParams>(w);
}

__device__ __forceinline__ void stw(uint32_t* p, const Fr& v) {
    uint32_t o[8];
    pack(canon(v), o);
    reinterpret_cast<uint4*>(p)[0] = make_uint4(o[0], o[1], o[2], o[3]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(o[4], o[5], o[6], o[7]);
}

__device__ __forceinline__ void stzero(uint32_t* p) {
    reinterpret_cast<uint4*>(p)[0] = make_uint4(0, 0, 0, 0);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(0, 0, 0, 0);
}

__device__ __forceinline__ void copyw(uint32_t* d, const uint32_t* s) {
    reinterpret_cast<uint4*>(d)[0] = reinterpret_cast<const uint4*>(s)[0];
    reinterpret_cast<uint4*>(d)[1] = reinterpret_cast<const uint4*>(s)[1];
}

// Words are read as residues (word value w stands for w R^-1): sums and
// differences need no correction, a product of two needs one factor R^2.
__device__ __forceinline__ Fr fix_product(const Fr& v) { return mul(v, fp_const<FrParams>(FrParams::R2)); }

// The inverse of the value whose word is w: as a Montgomery number w R
// (fix_product), inverted (w^-1 R), read back as a word (times R^-1).
__device__ __forceinline__ Fr word_inverse(const Fr& w) {
    const uint32_t one[8] = {1, 0, 0, 0, 0, 0, 0, 0};
    return mul(inv(fix_product(w)), unpack<FrParams>(one));
}

__host__ __device__ __forceinline__ uint64_t node_deg(uint64_t k, int l, uint64_t i) {
    const uint64_t lo = i << l, m = uint64_t(1) << l;
    if (lo >= k) return 0;
    return k - lo < m ? k - lo : m;
}

// Row r of a batched product reads operand row map(r) = (r % mod) * mul + add.
struct RowMap {
    uint64_t mod, mul, add;
    __device__ __forceinline__ uint64_t operator()(uint64_t r) const { return (r % mod) * mul + add; }
};
constexpr RowMap SAME_ROW = {~uint64_t(0), 1, 0};

// X[r] <- X[r] * H[m1(r)] (+ X2[r] * H2[m2(r)]), cyclic, length 2^L <= 32,
// in place: a block stages its rows in LDS first.
template <bool TWO>
__global__ void __launch_bounds__(BLOCK) k_conv_small(uint32_t* __restrict__ x, const uint32_t* __restrict__ h,
                                                      const uint32_t* __restrict__ x2,
                                                      const uint32_t* __restrict__ h2, uint64_t total, int L,
                                                      RowMap m1, RowMap m2) {
    __shared__ __align__(16) uint32_t sx[BLOCK * 8];
    __shared__ __align__(16) uint32_t sx2[TWO ? BLOCK * 8 : 8];
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    const uint32_t n = 1u << L;
    if (e < total) {
        copyw(sx + threadIdx.x * 8, x + e * 8);
        if (TWO) copyw(sx2 + threadIdx.x * 8, x2 + e * 8);
    }
    __syncthreads();
    if (e >= total) return;
    const uint64_t r = e >> L;
    const uint32_t j = uint32_t(e & (n - 1));
    const uint32_t base = threadIdx.x - j;  // this row's first element in LDS (blocks hold whole rows)
    const uint32_t* hr = h + m1(r) * n * 8;
    Fr acc = fp_zero<FrParams>();
    for (uint32_t i = 0; i < n; ++i)
        acc = add(acc, mul(ldw(sx + (base + i) * 8), ldw(hr + ((j - i) & (n - 1)) * 8)));
    if (TWO) {
        const uint32_t* hr2 = h2 + m2(r) * n * 8;
        for (uint32_t i = 0; i < n; ++i)
            acc = add(acc, mul(ldw(sx2 + (base + i) * 8), ldw(hr2 + ((j - i) & (n - 1)) * 8)));
    }
    stw(x + e * 8, fix_product(acc));
}

// Pointwise in the transformed domain: X[r][t] <- X[r][t] H[m1(r)][t] (+ X2[r][t] H2[m2(r)][t]).
template <bool TWO>
__global__ void __launch_bounds__(BLOCK) k_pointwise_rows(uint32_t* __restrict__ x, const uint32_t* __restrict__ h,
                                                          const uint32_t* __restrict__ x2,
                                                          const uint32_t* __restrict__ h2, uint64_t total, int L,
                                                          RowMap m1, RowMap m2) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t r = e >> L, t = e & ((uint64_t(1) << L) - 1);
    Fr acc = mul(ldw(x + e * 8), ldw(h + ((m1(r) << L) + t) * 8));
    if (TWO) acc = add(acc, mul(ldw(x2 + e * 8), ldw(h2 + ((m2(r) << L) + t) * 8)));
    stw(x + e * 8, fix_product(acc));
}

// ---- tree kernels (one thread per output element) ---------------------------

// lev0[i] = -x_i (i < k), 0 above.
__global__ void k_leaves(const uint32_t* __restrict__ pts, uint32_t* __restrict__ lev0, uint64_t k, uint64_t K) {
    const uint64_t i = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= K) return;
    if (i < k) stw(lev0 + i * 8, neg(ldw(pts + i * 8)));
    else stzero(lev0 + i * 8);
}

// Pair p of level l: X[p] = [A, 0], Y[p] = [B, 0] (rows of 2^(l+1)).
__global__ void k_pair_gather(const uint32_t* __restrict__ src, uint32_t* __restrict__ x, uint32_t* __restrict__ y,
                              uint64_t K, int l) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= K) return;
    const uint64_t m = uint64_t(1) << l, p = e >> (l + 1), t = e & (2 * m - 1);
    if (t < m) {
        copyw(x + e * 8, src + (2 * p * m + t) * 8);
        copyw(y + e * 8, src + ((2 * p + 1) * m + t) * 8);
    } else {
        stzero(x + e * 8);
        stzero(y + e * 8);
    }
}

// parent[t] = X[p][t] + (t >= a ? B[t - a]) + (t >= b ? A[t - b]) for t < a + b:
// the non-leading coefficients of (x^a + A)(x^b + B) with X = A B, and of
// P_A (x^b + B) + P_B (x^a + A) with X = P_A B + P_B A.
__global__ void k_pair_fix(const uint32_t* __restrict__ src, const uint32_t* __restrict__ x,
                           uint32_t* __restrict__ dst, uint64_t k, uint64_t K, int l) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= K) return;
    const uint64_t m = uint64_t(1) << l, p = e >> (l + 1), t = e & (2 * m - 1);
    const uint64_t a = node_deg(k, l, 2 * p), b = node_deg(k, l, 2 * p + 1);
    if (t >= a + b) {
        stzero(dst + e * 8);
        return;
    }
    Fr v = ldw(x + e * 8);
    if (t >= a) v = add(v, ldw(src + ((2 * p + 1) * m + t - a) * 8));
    if (t >= b) v = add(v, ldw(src + (2 * p * m + t - b) * 8));
    stw(dst + e * 8, v);
}

// Newton rows of one step: row r is node (l, i) of the levels l0..l1, the
// rows of level l0 first.
struct NewtonRows {
    uint64_t first[24];
    int l0, l1;
    __device__ __forceinline__ void find(uint64_t r, int& l, uint64_t& i) const {
        int j = l1 - l0;
        while (j > 0 && first[j] > r) --j;
        l = l0 + j;
        i = r - first[j];
    }
};

__global__ void k_rh_init(uint32_t* __restrict__ rh, uint64_t total, uint64_t K) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t l = e / (2 * K), t = (e % (2 * K)) & ((uint64_t(2) << l) - 1);
    if (t == 0) {
        reinterpret_cast<uint4*>(rh + e * 8)[0] = make_uint4(1, 0, 0, 0);
        reinterpret_cast<uint4*>(rh + e * 8)[1] = make_uint4(0, 0, 0, 0);
    } else {
        stzero(rh + e * 8);
    }
}

// X[r] = rev(node)[0..2m), Y[r] = [inv_m, 0^m].  Words: rev(node)[0] = 1 is
// the canonical word 1 as the other inputs are canonical words.
__global__ void k_newton_gather(const uint32_t* __restrict__ lev, const uint32_t* __restrict__ rh,
                                uint32_t* __restrict__ x, uint32_t* __restrict__ y, uint64_t total, uint64_t k,
                                uint64_t K, uint64_t m, NewtonRows rows) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t r = e / (2 * m), t = e % (2 * m);
    int l;
    uint64_t i;
    rows.find(r, l, i);
    const uint64_t d = node_deg(k, l, i), M = uint64_t(1) << l;
    if (t == 0) {
        reinterpret_cast<uint4*>(x + e * 8)[0] = make_uint4(1, 0, 0, 0);
        reinterpret_cast<uint4*>(x + e * 8)[1] = make_uint4(0, 0, 0, 0);
    } else if (t <= d) {
        copyw(x + e * 8, lev + (uint64_t(l) * K + i * M + d - t) * 8);
    } else {
        stzero(x + e * 8);
    }
    if (t < m) copyw(y + e * 8, rh + (uint64_t(l) * 2 * K + i * 2 * M + t) * 8);
    else stzero(y + e * 8);
}

// Z[r] = [X[r][m..2m), 0^m]
__global__ void k_newton_mid(const uint32_t* __restrict__ x, uint32_t* __restrict__ z, uint64_t total, uint64_t m) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t t = e % (2 * m);
    if (t < m) copyw(z + e * 8, x + (e + m) * 8);
    else stzero(z + e * 8);
}

// inv[m + t] = -Z[r][t], t < m
__global__ void k_newton_put(const uint32_t* __restrict__ z, uint32_t* __restrict__ rh, uint64_t total, uint64_t K,
                             uint64_t m, NewtonRows rows) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t r = e / m, t = e % m;
    int l;
    uint64_t i;
    rows.find(r, l, i);
    const uint64_t M = uint64_t(1) << l;
    stw(rh + (uint64_t(l) * 2 * K + i * 2 * M + m + t) * 8, neg(ldw(z + (r * 2 * m + t) * 8)));
}

// nh rows of every level: [node's coefficients, 0^(2^l)]
__global__ void k_nh_fill(const uint32_t* __restrict__ lev, uint32_t* __restrict__ nh, uint64_t total, uint64_t K) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t l = e / (2 * K), o = e % (2 * K), M = uint64_t(1) << l, i = o >> (l + 1), t = o & (2 * M - 1);
    if (t < M) copyw(nh + e * 8, lev + (l * K + i * M + t) * 8);
    else stzero(nh + e * 8);
}

// Remainder step from level l + 1 to level l, P polynomials, row r = (q, i):
// U[r][t] = r_parent[n - 1 - t] for t < qlen = d(sibling), n = d(parent).
__global__ void k_rem_gather(const uint32_t* __restrict__ rin, uint32_t* __restrict__ u, uint64_t total, uint64_t k,
                             uint64_t K, int l) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t M = uint64_t(1) << l, nodes = K >> l, r = e >> (l + 1), t = e & (2 * M - 1);
    const uint64_t q = r / nodes, i = r % nodes, p = i >> 1;
    const uint64_t dc = node_deg(k, l, i), ds = node_deg(k, l, i ^ 1), n = dc + ds;
    if (t < ds) copyw(u + e * 8, rin + (q * K + p * 2 * M + n - 1 - t) * 8);
    else stzero(u + e * 8);
}

// V[r][t] = U[r][qlen - 1 - t] for t < qlen: the quotient
__global__ void k_rem_reverse(const uint32_t* __restrict__ u, uint32_t* __restrict__ v, uint64_t total, uint64_t k,
                              uint64_t K, int l) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t M = uint64_t(1) << l, nodes = K >> l, r = e >> (l + 1), t = e & (2 * M - 1);
    const uint64_t ql = node_deg(k, l, (r % nodes) ^ 1);
    if (t < ql) copyw(v + e * 8, u + ((r << (l + 1)) + ql - 1 - t) * 8);
    else stzero(v + e * 8);
}

// r_child[t] = r_parent[t] - (q C)[t], t < d(child); zeros to 2^l
__global__ void k_rem_put(const uint32_t* __restrict__ rin, const uint32_t* __restrict__ v,
                          uint32_t* __restrict__ rout, uint64_t total, uint64_t k, uint64_t K, int l) {
    const uint64_t e = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= total) return;
    const uint64_t M = uint64_t(1) << l, nodes = K >> l, r = e >> l, t = e & (M - 1);
    const uint64_t q = r / nodes, i = r % nodes, p = i >> 1;
    if (t < node_deg(k, l, i))
        stw(rout + e * 8, sub(ldw(rin + (q * K + p * 2 * M + t) * 8), ldw(v + ((r << (l + 1)) + t) * 8)));
    else
        stzero(rout + e * 8);
}

// dst[t] = (t + 1) src[t + 1] for t + 1 < n, plus the leading term's
// derivative: monic A of degree n has src = its n low coefficients and
// dst[n - 1] = n.  K outputs.
__global__ void k_diff(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint64_t n, uint64_t K,
                       int monic) {
    const uint64_t t = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (t >= K) return;
    uint32_t w[8] = {uint32_t(t + 1), uint32_t((t + 1) >> 32), 0, 0, 0, 0, 0, 0};
    if (t + 1 < n) {
        stw(dst + t * 8, fix_product(mul(ldw(src + (t + 1) * 8), unpack<FrParams>(w))));
    } else if (monic && t + 1 == n) {
        stw(dst + t * 8, unpack<FrParams>(w));
    } else {
        stzero(dst + t * 8);
    }
}

// ideriv[i] = 1 / deriv[i] (Fermat); flag when some A'(x_i), i < k, is zero.
__global__ void k_invert(const uint32_t* __restrict__ d, uint32_t* __restrict__ out, uint64_t k, uint64_t K,
                         uint32_t* __restrict__ flag) {
    const uint64_t i = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= K) return;
    if (i >= k) {
        stzero(out + i * 8);
        return;
    }
    const Fr v = ldw(d + i * 8);
    if (is_zero(v)) {
        atomicOr(flag, 1u);
        stzero(out + i * 8);
        return;
    }
    stw(out + i * 8, word_inverse(v));
}

// out[i] = a[i] b[i] (i < k), 0 above
__global__ void k_mul_elems(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t* __restrict__ out,
                            uint64_t k, uint64_t K) {
    const uint64_t i = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= K) return;
    if (i < k) stw(out + i * 8, fix_product(mul(ldw(a + i * 8), ldw(b + i * 8))));
    else stzero(out + i * 8);
}

// fast.rs:118-139: (N'(x_i) - w_i A''(x_i) / 2) / A'(x_i)
__global__ void k_cauchy(const uint32_t* __restrict__ first, const uint32_t* __restrict__ second,
                         const uint32_t* __restrict__ w, const uint32_t* __restrict__ ideriv,
                         uint32_t* __restrict__ out, uint64_t k, Fr half) {
    const uint64_t i = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= k) return;
    const Fr ws = fix_product(mul(ldw(w + i * 8), ldw(second + i * 8)));
    const Fr hs = fix_product(mul(ws, half));
    stw(out + i * 8, fix_product(mul(sub(ldw(first + i * 8), hs), ldw(ideriv + i * 8))));
}

}  // namespace

struct DevTree {
    Owner* o = nullptr;
    uint64_t k = 0, K = 0;
    int L = 0;
    uint32_t root[8] = {0};
    DeviceBuffer lev, rh, nh, deriv, ideriv, x, y, z, ra, rb, flag;
    DevTreeStats* st = nullptr;

    uint32_t* P(DeviceBuffer& b, uint64_t elem = 0) { return static_cast<uint32_t*>(b.ptr) + elem * 8; }
    hipStream_t stream() const { return o->stream; }

    void domain() { ntt_domain_prepare(o->domain, root, o->stream); }

    // Transform `rows` operand rows of 2^L2 elements in place (nothing for
    // the schoolbook lengths, whose operands stay coefficients).
    void prep(uint32_t* h, uint64_t rows, int L2) {
        if (L2 > SMALL_LOG) ntt_rows(o->domain, h, h, L2, rows, false, o->ntt, stream());
    }

    // X <- X * H (+ X2 * H2), cyclic length 2^L2, H/H2 prepared.
    void mul_pre(uint32_t* xs, const uint32_t* h, uint32_t* x2, const uint32_t* h2, uint64_t rows, int L2,
                 RowMap m1, RowMap m2) {
        const uint64_t total = rows << L2;
        const unsigned grid = grid_for(total, BLOCK);
        const bool two = x2 != nullptr;
        if (L2 <= SMALL_LOG) {
            if (two) k_conv_small<true><<<grid, BLOCK, 0, stream()>>>(xs, h, x2, h2, total, L2, m1, m2);
            else k_conv_small<false><<<grid, BLOCK, 0, stream()>>>(xs, h, xs, h, total, L2, m1, m2);
            HIP_CHECK_LAUNCH();
        } else {
            // a second operand stored right after the first: one transform call for both
            const bool adjacent = two && x2 == xs + (rows << L2) * 8;
            ntt_rows(o->domain, xs, xs, L2, adjacent ? 2 * rows : rows, false, o->ntt, stream());
            if (two && !adjacent) ntt_rows(o->domain, x2, x2, L2, rows, false, o->ntt, stream());
            if (two) k_pointwise_rows<true><<<grid, BLOCK, 0, stream()>>>(xs, h, x2, h2, total, L2, m1, m2);
            else k_pointwise_rows<false><<<grid, BLOCK, 0, stream()>>>(xs, h, xs, h, total, L2, m1, m2);
            HIP_CHECK_LAUNCH();
            ntt_rows(o->domain, xs, xs, L2, rows, true, o->ntt, stream());
        }
        ++st->rounds;
        st->products += two ? 2 * rows : rows;
    }

    // X <- X * Y row by row (Y is left transformed).  Y stored right after X
    // is transformed in the same call as X.
    void mul(uint32_t* xs, uint32_t* ys, uint64_t rows, int L2) {
        const uint64_t total = rows << L2;
        if (L2 > SMALL_LOG && ys == xs + total * 8) {
            ntt_rows(o->domain, xs, xs, L2, 2 * rows, false, o->ntt, stream());
            const unsigned grid = grid_for(total, BLOCK);
            k_pointwise_rows<false><<<grid, BLOCK, 0, stream()>>>(xs, ys, xs, ys, total, L2, SAME_ROW, SAME_ROW);
            HIP_CHECK_LAUNCH();
            ntt_rows(o->domain, xs, xs, L2, rows, true, o->ntt, stream());
            ++st->rounds;
            st->products += rows;
            return;
        }
        prep(ys, rows, L2);
        mul_pre(xs, ys, nullptr, nullptr, rows, L2, SAME_ROW, SAME_ROW);
    }

    // Values at the leaves of `count` polynomials placed at ra[q K ...].
    // Leaves the values in ra[q K + i] or rb (returned).
    uint32_t* remainder_tree(uint64_t count) {
        uint32_t *rin = P(ra), *rout = P(rb);
        for (int l = L - 1; l >= 0; --l) {
            const uint64_t nodes = K >> l, rows = count * nodes, total = rows << (l + 1);
            const RowMap by_node = {nodes, 1, 0};
            k_rem_gather<<<grid_for(total, BLOCK), BLOCK, 0, stream()>>>(rin, P(x), total, k, K, l);
            HIP_CHECK_LAUNCH();
            mul_pre(P(x), P(rh, uint64_t(l) * 2 * K), nullptr, nullptr, rows, l + 1, by_node, by_node);
            k_rem_reverse<<<grid_for(total, BLOCK), BLOCK, 0, stream()>>>(P(x), P(y), total, k, K, l);
            HIP_CHECK_LAUNCH();
            mul_pre(P(y), P(nh, uint64_t(l) * 2 * K), nullptr, nullptr, rows, l + 1, by_node, by_node);
            const uint64_t outs = rows << l;
            k_rem_put<<<grid_for(outs, BLOCK), BLOCK, 0, stream()>>>(rin, P(y), rout, outs, k, K, l);
            HIP_CHECK_LAUNCH();
            std::swap(rin, rout);
        }
        return rin;
    }

    // Interpolation upward from leaf constants in ra (K elements); returns
    // the buffer holding the root's coefficients.
    uint32_t* interpolation_tree() {
        uint32_t *vin = P(ra), *vout = P(rb);
        for (int l = 0; l < L; ++l) {
            const uint64_t pairs = K >> (l + 1);
            k_pair_gather<<<grid_for(K, BLOCK), BLOCK, 0, stream()>>>(vin, P(x), P(x, K), K, l);
            HIP_CHECK_LAUNCH();
            // P_A B + P_B A: right child's row with P_A, left child's with P_B
            mul_pre(P(x), P(nh, uint64_t(l) * 2 * K), P(x, K), P(nh, uint64_t(l) * 2 * K), pairs, l + 1,
                    RowMap{~uint64_t(0), 2, 1}, RowMap{~uint64_t(0), 2, 0});
            k_pair_fix<<<grid_for(K, BLOCK), BLOCK, 0, stream()>>>(vin, P(x), vout, k, K, l);
            HIP_CHECK_LAUNCH();
            std::swap(vin, vout);
        }
        return vin;
    }

    void upload(uint32_t* dst, const uint32_t* src, uint64_t n, uint64_t pad_to) {
        if (n) HIP_CHECK(hipMemcpyAsync(dst, src, n * 32, hipMemcpyHostToDevice, stream()));
        if (pad_to > n) HIP_CHECK(hipMemsetAsync(dst + n * 8, 0, (pad_to - n) * 32, stream()));
    }

    void download(uint32_t* dst, const uint32_t* src, uint64_t n) {
        if (n) HIP_CHECK(hipMemcpyAsync(dst, src, n * 32, hipMemcpyDeviceToHost, stream()));
        HIP_CHECK(hipStreamSynchronize(stream()));
    }
};

namespace {

struct StatsScope {
    DevTree* t;
    StatsScope(DevTree* tree, DevTreeStats& s) : t(tree) { t->st = &s; }
    ~StatsScope() { t->st = nullptr; }
};

}  // namespace

DevTree* dev_tree_build(Owner* o, const uint32_t* points, size_t k, DevTreeStats& st) {
    require(k >= 2 && k <= DEV_TREE_MAX_POINTS, "device tree point count out of range");
    require(o->domain.ready, "NTT domain not prepared");
    auto t = std::make_unique<DevTree>();
    t->o = o;
    t->k = k;
    t->K = 1;
    while (t->K < k) {
        t->K <<= 1;
        ++t->L;
    }
    std::memcpy(t->root, o->domain.root_words, 32);
    StatsScope scope(t.get(), st);
    const uint64_t K = t->K;
    const int L = t->L;
    hipStream_t s = o->stream;
    t->lev.ensure(uint64_t(L + 1) * K * 32);
    t->rh.ensure(uint64_t(L) * 2 * K * 32);
    t->nh.ensure(uint64_t(L) * 2 * K * 32);
    t->deriv.ensure(K * 32);
    t->ideriv.ensure(K * 32);
    for (DeviceBuffer* b : {&t->x, &t->y, &t->z}) b->ensure(4 * K * 32);
    t->ra.ensure(2 * K * 32);
    t->rb.ensure(2 * K * 32);
    t->flag.ensure(256);
    // levels
    t->upload(t->P(t->x), points, k, 0);
    k_leaves<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(t->P(t->x), t->P(t->lev), k, K);
    HIP_CHECK_LAUNCH();
    for (int l = 0; l < L; ++l) {
        const uint64_t pairs = K >> (l + 1);
        uint32_t* src = t->P(t->lev, uint64_t(l) * K);
        k_pair_gather<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(src, t->P(t->x), t->P(t->x, K), K, l);
        HIP_CHECK_LAUNCH();
        t->mul(t->P(t->x), t->P(t->x, K), pairs, l + 1);
        k_pair_fix<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(src, t->P(t->x), t->P(t->lev, uint64_t(l + 1) * K), k, K,
                                                         l);
        HIP_CHECK_LAUNCH();
    }
    // reciprocals rev(node)^-1 mod x^(2^l) for the levels 1..L-1 (level 0: 1)
    const uint64_t rh_total = uint64_t(L) * 2 * K;
    k_rh_init<<<grid_for(rh_total, BLOCK), BLOCK, 0, s>>>(t->P(t->rh), rh_total, K);
    HIP_CHECK_LAUNCH();
    for (int sbit = 0; sbit + 2 <= L; ++sbit) {  // m = 2^sbit; levels sbit + 1 .. L - 1
        const uint64_t m = uint64_t(1) << sbit;
        NewtonRows rows{};
        rows.l0 = sbit + 1;
        rows.l1 = L - 1;
        uint64_t count = 0;
        for (int l = rows.l0; l <= rows.l1; ++l) {
            rows.first[l - rows.l0] = count;
            count += K >> l;
        }
        const uint64_t total = count * 2 * m;
        uint32_t *xs = t->P(t->x), *ys = t->P(t->x, total);  // Y right after X (total <= 2K)
        k_newton_gather<<<grid_for(total, BLOCK), BLOCK, 0, s>>>(t->P(t->lev), t->P(t->rh), xs, ys, total, k, K, m,
                                                                 rows);
        HIP_CHECK_LAUNCH();
        t->mul(xs, ys, count, sbit + 1);
        k_newton_mid<<<grid_for(total, BLOCK), BLOCK, 0, s>>>(xs, t->P(t->z), total, m);
        HIP_CHECK_LAUNCH();
        t->mul_pre(t->P(t->z), ys, nullptr, nullptr, count, sbit + 1, SAME_ROW, SAME_ROW);
        const uint64_t puts = count * m;
        k_newton_put<<<grid_for(puts, BLOCK), BLOCK, 0, s>>>(t->P(t->z), t->P(t->rh), puts, K, m, rows);
        HIP_CHECK_LAUNCH();
    }
    k_nh_fill<<<grid_for(rh_total, BLOCK), BLOCK, 0, s>>>(t->P(t->lev), t->P(t->nh), rh_total, K);
    HIP_CHECK_LAUNCH();
    for (int l = 0; l < L; ++l) {
        t->prep(t->P(t->rh, uint64_t(l) * 2 * K), K >> l, l + 1);
        t->prep(t->P(t->nh, uint64_t(l) * 2 * K), K >> l, l + 1);
    }
    // A'(x_i) by the remainder tree, then 1 / A'(x_i)
    k_diff<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(t->P(t->lev, uint64_t(L) * K), t->P(t->ra), k, K, 1);
    HIP_CHECK_LAUNCH();
    uint32_t* vals = t->remainder_tree(1);
    HIP_CHECK(hipMemcpyAsync(t->P(t->deriv), vals, K * 32, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemsetAsync(t->flag.ptr, 0, 4, s));
    k_invert<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(t->P(t->deriv), t->P(t->ideriv), k, K,
                                                  static_cast<uint32_t*>(t->flag.ptr));
    HIP_CHECK_LAUNCH();
    uint32_t flag = 0;
    HIP_CHECK(hipMemcpyAsync(&flag, t->flag.ptr, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    require(flag == 0, "repeated interpolation root");
    return t.release();
}

void dev_tree_destroy(DevTree* t) noexcept { delete t; }

size_t dev_tree_points(const DevTree* t) { return t->k; }

void dev_tree_vanishing(DevTree* t, uint32_t* out) {
    t->download(out, t->P(t->lev, uint64_t(t->L) * t->K), t->k);
    std::memset(out + t->k * 8, 0, 32);
    out[t->k * 8] = 1;
}

void dev_tree_inverse_derivatives(DevTree* t, uint32_t* out) { t->download(out, t->P(t->ideriv), t->k); }

void dev_tree_eval(DevTree* t, const uint32_t* const* polys, const size_t* lens, size_t count,
                   uint32_t* const* values, DevTreeStats& st) {
    require(count >= 1 && count <= 2, "device tree evaluates one or two polynomials a call");
    StatsScope scope(t, st);
    t->domain();
    for (size_t q = 0; q < count; ++q) {
        require(lens[q] <= t->k, "polynomial longer than the tree's point count");
        t->upload(t->P(t->ra, q * t->K), polys[q], lens[q], t->K);
    }
    uint32_t* vals = t->remainder_tree(count);
    for (size_t q = 0; q < count; ++q)
        HIP_CHECK(hipMemcpyAsync(values[q], vals + q * t->K * 8, t->k * 32, hipMemcpyDeviceToHost, t->stream()));
    HIP_CHECK(hipStreamSynchronize(t->stream()));
}

void dev_tree_interpolate(DevTree* t, const uint32_t* values, uint32_t* coeffs, DevTreeStats& st) {
    StatsScope scope(t, st);
    t->domain();
    t->upload(t->P(t->z), values, t->k, 0);
    k_mul_elems<<<grid_for(t->K, BLOCK), BLOCK, 0, t->stream()>>>(t->P(t->z), t->P(t->ideriv), t->P(t->ra), t->k,
                                                                    t->K);
    HIP_CHECK_LAUNCH();
    t->download(coeffs, t->interpolation_tree(), t->k);
}

void dev_tree_cauchy(DevTree* t, const uint32_t* weights, uint32_t* sums, DevTreeStats& st) {
    StatsScope scope(t, st);
    t->domain();
    const uint64_t K = t->K, k = t->k;
    hipStream_t s = t->stream();
    // N = the interpolant of w_j A'(x_j): its leaf constants w_j A'(x_j) / A'(x_j) are w_j
    t->upload(t->P(t->ra), weights, k, K);
    uint32_t* num = t->interpolation_tree();
    // d1 = N' and d2 = A'' side by side in ra (N copied out of ra/rb first,
    // A' formed in x)
    HIP_CHECK(hipMemcpyAsync(t->P(t->z), num, K * 32, hipMemcpyDeviceToDevice, s));
    k_diff<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(t->P(t->z), t->P(t->ra), k, K, 0);
    HIP_CHECK_LAUNCH();
    k_diff<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(t->P(t->lev, uint64_t(t->L) * K), t->P(t->x), k, K, 1);
    HIP_CHECK_LAUNCH();
    k_diff<<<grid_for(K, BLOCK), BLOCK, 0, s>>>(t->P(t->x), t->P(t->ra, K), k, K, 0);
    HIP_CHECK_LAUNCH();
    uint32_t* vals = t->remainder_tree(2);
    // the weights again (ra and rb were overwritten), then the combination
    t->upload(t->P(t->z), weights, k, 0);
    // the word of 1/2 = (r + 1) / 2 = (r >> 1) + 1 (r odd; k_cauchy's operands are words)
    uint32_t hw[8];
    for (int q = 0; q < 8; ++q) hw[q] = (FR_MOD32[q] >> 1) | (q < 7 ? FR_MOD32[q + 1] << 31 : 0u);
    hw[0] += 1;  // r >> 1 ends in ...f8000000: no carry
    k_cauchy<<<grid_for(k, BLOCK), BLOCK, 0, s>>>(vals, vals + K * 8, t->P(t->z), t->P(t->ideriv), t->P(t->y), k,
                                                  unpack<FrParams>(hw));
    HIP_CHECK_LAUNCH();
    t->download(sums, t->P(t->y), k);
}

}  // namespace dvgpu
