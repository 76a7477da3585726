// Pippenger MSM for BN254 G1/G2 on gfx950.
//
// Phases (one stream, no host synchronisation until the window values):
//   1. digits   : scalar -> ceil(255/c) signed c-bit digits; each nonzero digit
//                 takes a rank in its bucket with one global atomic.
//   (Every kernel serves a whole batch of MSMs: a thread finds its MSM by a
//   binary search over the descriptor table, whose prefix bases place each
//   MSM's entries, buckets, chunks and windows in global arrays.)
//   2. scan     : exclusive scan of (bucket count, task count) pairs.
//   3. scatter  : point index (plus sign bit) to its bucket-sorted slot.
//   4. tasks    : every bucket is cut into tasks of at most T points, so one
//                 hot bucket (repeated scalars) spreads over many threads.
//   5. accumulate: one thread per task, XYZZ mixed additions -> partial sums.
//   6. reduce   : one thread per chunk of L buckets computes
//                 sum_i (i+1) S_i by running sums and adds (jL) * sum_i S_i.
//   7. tree     : one block per window sums its chunk values (LDS tree).
// The W window values go to the host, which forms sum_w 2^(wc) V_w: ~250
// serial doublings that one CPU core finishes far sooner than one GPU lane.
#pragma once
#include "msm.hpp"

#include <algorithm>
#include <atomic>
#include <memory>
#include <thread>
#include <cstring>
#include <type_traits>
#include <utility>

#include "host_g1.hpp"
#include "host_g2.hpp"
#include "host_pool.hpp"

namespace dvgpu {

using namespace bn254;

// ---------------------------------------------------------------------------
// Point memory access
// ---------------------------------------------------------------------------

template <class F>
struct Mem;

template <>
struct Mem<FqOps> {
    static __device__ __forceinline__ Fq ld(const uint32_t* p) {
        const uint4 a = reinterpret_cast<const uint4*>(p)[0];
        const uint4 b = reinterpret_cast<const uint4*>(p)[1];
        const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        return unpack<FqParams>(w);
    }
    static __device__ __forceinline__ void st(uint32_t* p, const Fq& v) {
        uint32_t w[8];
        pack(v, w);
        reinterpret_cast<uint4*>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
        reinterpret_cast<uint4*>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
    static __device__ __forceinline__ Affine<FqOps> load_affine(const uint32_t* p) {
        const uint4* q = reinterpret_cast<const uint4*>(p);
        const uint4 a = q[0], b = q[1], c = q[2], d = q[3];
        const uint32_t x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const uint32_t y[8] = {c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
        const uint32_t any = a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w | c.x | c.y | c.z |
                             c.w | d.x | d.y | d.z | d.w;
        return {unpack<FqParams>(x), unpack<FqParams>(y), any == 0};
    }
    static __device__ __forceinline__ Xyzz<FqOps> load_xyzz(const uint32_t* p) {
        return {ld(p), ld(p + 8), ld(p + 16), ld(p + 24)};
    }
    static __device__ __forceinline__ void store_xyzz(uint32_t* p, const Xyzz<FqOps>& v) {
        st(p, v.X);
        st(p + 8, v.Y);
        st(p + 16, v.ZZ);
        st(p + 24, v.ZZZ);
    }
};

template <>
struct Mem<Fq2Ops> {
    static __device__ __forceinline__ Fq2 ld(const uint32_t* p) {
        return {Mem<FqOps>::ld(p), Mem<FqOps>::ld(p + 8)};
    }
    static __device__ __forceinline__ void st(uint32_t* p, const Fq2& v) {
        Mem<FqOps>::st(p, v.c0);
        Mem<FqOps>::st(p + 8, v.c1);
    }
    static __device__ __forceinline__ Affine<Fq2Ops> load_affine(const uint32_t* p) {
        uint32_t any = 0;
        const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint4 v = q[i];
            any |= v.x | v.y | v.z | v.w;
        }
        return {ld(p), ld(p + 16), any == 0};
    }
    static __device__ __forceinline__ Xyzz<Fq2Ops> load_xyzz(const uint32_t* p) {
        return {ld(p), ld(p + 16), ld(p + 32), ld(p + 48)};
    }
    static __device__ __forceinline__ void store_xyzz(uint32_t* p, const Xyzz<Fq2Ops>& v) {
        st(p, v.X);
        st(p + 16, v.Y);
        st(p + 32, v.ZZ);
        st(p + 48, v.ZZZ);
    }
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

constexpr uint32_t NO_DIGIT = 0xffffffffu;

// Width and lowest bit of window w (MsmDesc::wide).
__host__ __device__ __forceinline__ uint32_t window_width(const MsmDesc& d, uint32_t w) {
    return w < d.wide ? d.c : d.c - 1;
}
__host__ __device__ __forceinline__ uint32_t window_offset(const MsmDesc& d, uint32_t w) {
    return w <= d.wide ? w * d.c : d.wide * d.c + (w - d.wide) * (d.c - 1);
}

// Full-width windows of W windows of c or c - 1 bits covering 255 bits.
__host__ __device__ __forceinline__ uint32_t wide_windows(uint32_t c, uint32_t W) {
    const int64_t wide = 255 - int64_t(W) * int64_t(c - 1);
    return uint32_t(wide < 0 ? 0 : (wide > int64_t(W) ? int64_t(W) : wide));
}

// Accumulation tasks: every bucket of c points is cut into floor(c / T)
// full tasks of T points and, when T does not divide c, one task of c mod T
// points.  Tasks run longest first (class 0: the full tasks, class k: tasks
// of T - k points), so that the 64 lanes of a wave add about equally many
// points; in bucket order a wave's lanes mix full tasks with short
// remainders and all wait for the longest.
constexpr uint32_t TASK_CLASS_MAX = 1024;  // task_points <= 1024
// Buckets cut into more than HEAVY_TASKS tasks have their partial sums
// combined by a whole block (k_heavy); the others with two or more tasks by
// one thread (k_bucket_sums), which takes them in order of decreasing task
// count (bucket-sum class HEAVY_TASKS - tasks) so that a wave's lanes chain
// about equally many additions.
constexpr uint32_t HEAVY_TASKS = 32;
// Words of the workspace's counter area: [0] entries, [1] tasks, [2] heavy
// buckets, [3] error word, then the bucket-sum class counts and cursors, the
// number of buckets listed for k_bucket_sums, and the task-class counts and
// cursors.
constexpr uint32_t TOTAL_SUM_COUNTS = 4;
constexpr uint32_t TOTAL_SUM_CURSORS = TOTAL_SUM_COUNTS + HEAVY_TASKS;
constexpr uint32_t TOTAL_SUM_LISTED = TOTAL_SUM_CURSORS + HEAVY_TASKS;
constexpr uint32_t TOTAL_CLASS_COUNTS = TOTAL_SUM_LISTED + 4;
constexpr uint32_t TOTAL_CLASS_CURSORS = TOTAL_CLASS_COUNTS + TASK_CLASS_MAX;
constexpr uint32_t TOTAL_WORDS = TOTAL_CLASS_CURSORS + TASK_CLASS_MAX;

// Bucket-sum class of a bucket of `count` points, or HEAVY_TASKS when no
// thread sums it (fewer than two tasks, or a heavy bucket).
__host__ __device__ __forceinline__ uint32_t sum_class(uint32_t count, uint32_t tp) {
    const uint32_t nt = (count + tp - 1) / tp;
    return nt >= 2 && nt <= HEAVY_TASKS ? HEAVY_TASKS - nt : HEAVY_TASKS;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Adds the task classes of `count` buckets (one per thread, 0 for none) to
// the LDS histogram h (tp zeroed counters; full tasks summed per wave) and
// their bucket-sum classes to hs (HEAVY_TASKS zeroed counters).
__device__ __forceinline__ void task_class_add(uint32_t count, uint32_t tp, uint32_t* h, uint32_t* hs) {
    const uint32_t r = count % tp;
    if (r) atomicAdd(&h[tp - r], 1u);
    const uint32_t full = wave_sum(count / tp);
    if ((threadIdx.x & 63) == 0 && full) atomicAdd(&h[0], full);
    const uint32_t sc = sum_class(count, tp);
    if (sc < HEAVY_TASKS) atomicAdd(&hs[sc], 1u);
}

// The block's class histograms added to the global counts.
__device__ __forceinline__ void task_class_flush(uint32_t tp, const uint32_t* h, const uint32_t* hs,
                                                 uint32_t* __restrict__ total) {
    for (uint32_t i = threadIdx.x; i < tp; i += blockDim.x)
        if (h[i]) atomicAdd(&total[TOTAL_CLASS_COUNTS + i], h[i]);
    if (threadIdx.x < HEAVY_TASKS && hs[threadIdx.x]) atomicAdd(&total[TOTAL_SUM_COUNTS + threadIdx.x], hs[threadIdx.x]);
}

// bits [bit, bit + c) of a 256-bit little-endian scalar (c <= 24)
// Signed digits in [-(2^(b-1) - 1), 2^(b-1)] (b the window's width) by one
// addition: with K = sum_k (2^(b_k-1) - 1) 2^(o_k) (o_k the window's lowest
// bit), digit k of s is window k of t = s + K minus (2^(b_k-1) - 1) (the
// representation is unique, so this equals the carry-propagating
// recoding).  t needs nine words once the windows cover more than 256 bits.
__device__ __forceinline__ void recode_constant(const MsmDesc& d, uint32_t (&k)[9]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) k[i] = 0;
    for (uint32_t w = 0; w < d.digit_windows; ++w) {
        const uint64_t low = (uint64_t(1) << (window_width(d, w) - 1)) - 1;
        const uint32_t bit = window_offset(d, w), q = bit >> 5, sh = bit & 31;
        const uint64_t v = low << sh;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            if (uint32_t(i) == q) k[i] |= uint32_t(v);
            if (uint32_t(i) == q + 1) k[i] |= uint32_t(v >> 32);
        }
    }
}

__device__ __forceinline__ void recode_add(const uint32_t (&s)[8], const uint32_t (&k)[9], uint32_t (&t)[9]) {
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const uint64_t x = uint64_t(i < 8 ? s[i] : 0u) + k[i] + carry;
        t[i] = uint32_t(x);
        carry = x >> 32;
    }
}

// Window `w` of the nine-word t (no dynamic register indexing: selects).
__device__ __forceinline__ uint32_t window_of(const uint32_t (&t)[9], uint32_t bit, int c) {
    const uint32_t q = bit >> 5, sh = bit & 31;
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (uint32_t(i) == q) lo = t[i];
        if (uint32_t(i) == q + 1) hi = t[i];
    }
    return uint32_t(((uint64_t(hi) << 32) | lo) >> sh) & ((1u << c) - 1u);
}

__device__ __forceinline__ uint32_t scalar_bits(const uint32_t (&s)[8], int bit, int c) {
    const int j = bit >> 5, sh = bit & 31;
    uint64_t v = 0;
    if (j < 8) v = s[j];
    if (j + 1 < 8) v |= uint64_t(s[j + 1]) << 32;
    return uint32_t(v >> sh) & ((1u << c) - 1u);
}


// Index of the descriptor owning `key` under the prefix field `field`: the
// last descriptor whose base is <= key.
template <uint32_t MsmDesc::*field>
__device__ __forceinline__ uint32_t find_desc(const MsmDesc* __restrict__ d, uint32_t m, uint32_t key) {
    uint32_t lo = 0, hi = m - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (d[mid].*field <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

static __global__ void k_fill(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

static __global__ void k_digits(const MsmDesc* __restrict__ descs, uint32_t m, uint32_t total_threads,
                                const uint32_t* __restrict__ scalars, uint32_t* __restrict__ counts,
                                uint32_t* __restrict__ digits, uint32_t* __restrict__ ranks) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total_threads) return;
    const MsmDesc d = descs[find_desc<&MsmDesc::thread_base>(descs, m, t)];
    const uint32_t i = t - d.thread_base;
    const uint4* sp = reinterpret_cast<const uint4*>(scalars + size_t(d.scalar_base + i) * 8);
    const uint4 a = sp[0], b = sp[1];
    const uint32_t s[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t carry = 0;
    for (uint32_t w = 0; w < d.digit_windows; ++w) {
        const int c = int(window_width(d, w));
        uint32_t v = scalar_bits(s, int(window_offset(d, w)), c) + carry;
        uint32_t sign = 0;
        if (v > (1u << (c - 1))) {
            v = (1u << c) - v;
            sign = 1;
            carry = 1;
        } else {
            carry = 0;
        }
        const size_t slot = size_t(d.entry_base) + size_t(w) * d.n + i;
        if (v == 0) {
            digits[slot] = NO_DIGIT;
        } else {
            const uint32_t bucket = v - 1;
            const uint32_t row = d.table ? 0u : w;  // table mode: one bucket set for all windows
            ranks[slot] = atomicAdd(&counts[d.bucket_base + row * d.buckets + bucket], 1u);
            digits[slot] = (bucket << 1) | sign;
        }
    }
}

// Digit planes (group mode): every scalar is read and recoded once, its
// signed digits written as 16-bit entries, plane w of MSM d at
// d.entry_base + w d.n (magnitude | sign << 15; 0 = no digit; magnitudes are
// at most 2^14 as bucket sets fit the LDS histogram).  The counting and
// staging blocks, one per (chunk, window), then read two bytes per point
// instead of the whole scalar.
static __global__ void __launch_bounds__(256)
    k_digit_planes(const MsmDesc* __restrict__ descs, uint32_t m, uint32_t total_threads,
                   const uint32_t* __restrict__ scalars, uint16_t* __restrict__ planes) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total_threads) return;
    const MsmDesc d = descs[find_desc<&MsmDesc::thread_base>(descs, m, t)];
    const uint32_t i = t - d.thread_base;
    const uint4* sp = reinterpret_cast<const uint4*>(scalars + size_t(d.scalar_base + i) * 8);
    const uint4 a = sp[0], b = sp[1];
    const uint32_t s[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t kconst[9], r[9];
    recode_constant(d, kconst);
    recode_add(s, kconst, r);
    for (uint32_t w = 0; w < d.digit_windows; ++w) {
        const int c = int(window_width(d, w));
        const int32_t digit = int32_t(window_of(r, window_offset(d, w), c)) - int32_t((1u << (c - 1)) - 1);
        const uint32_t v = uint32_t(digit < 0 ? -digit : digit), sign = digit < 0 ? 1u : 0u;
        planes[size_t(d.entry_base) + size_t(w) * d.n + i] = uint16_t(v | (sign << 15));
    }
}

// Blocks of the LDS digit path map to XCDs round-robin; consecutive logical
// blocks (the windows of one chunk of points) are placed on one XCD so that
// its L2 serves the chunk's scalars to all of them.
__device__ __forceinline__ uint32_t xcd_grouped_block(uint32_t bid, uint32_t grid) {
    constexpr uint32_t XCDS = 8;
    const uint32_t q = grid / XCDS, r = grid % XCDS, x = bid % XCDS, pos = bid / XCDS;
    return x < r ? x * (q + 1) + pos : r * (q + 1) + (x - r) * q + pos;
}

// One block per (MSM, chunk of points, window), run twice.  Counting pass:
// the signed digit of every point of the chunk in that window (the carry in
// is recomputed from the lower windows) is counted in an LDS histogram, which
// goes to global memory for the per-bucket prefix over blocks.  Placing pass:
// the block's start per bucket is the bucket's start (the scan's offsets)
// plus the block's prefix among the bucket's blocks; the block recomputes
// the digits and places every point with an LDS atomic on that start.  (The
// order within a bucket is arbitrary; a bucket's sum does not depend on it.)
// PLANES (the counting pass of group mode): the digits come from the
// planes of k_digit_planes instead of the scalars.
template <bool PLACE, bool PLANES = false>
static __global__ void __launch_bounds__(MSM_DIGIT_THREADS)
    k_digits_lds(const MsmDesc* __restrict__ descs, uint32_t m, uint32_t grid, const uint32_t* __restrict__ scalars,
                 uint32_t* __restrict__ hist, const uint2* __restrict__ offsets,
                 const uint32_t* __restrict__ point_index, uint32_t point_limit,
                 uint32_t* __restrict__ flags, uint32_t* __restrict__ sorted,
                 const uint16_t* __restrict__ planes = nullptr) {
    __shared__ uint32_t counts[MSM_LDS_BUCKETS];
    const uint32_t g = xcd_grouped_block(blockIdx.x, grid);
    const MsmDesc d = descs[find_desc<&MsmDesc::hist_block_base>(descs, m, g)];
    const uint32_t local = g - d.hist_block_base;
    // Normal mode: block = (chunk j, window w).  Table mode: every window
    // shares the bucket set, so one block covers all windows of its chunk.
    const uint32_t j = d.table ? local : local / d.digit_windows;
    const uint32_t w0 = d.table ? 0u : local % d.digit_windows, w1 = d.table ? d.digit_windows : w0 + 1;
    uint32_t* mine = hist + d.hist_base + size_t(local) * d.buckets;
    const uint2* starts = offsets + d.bucket_base + (d.table ? 0u : w0) * d.buckets;
    for (uint32_t b = threadIdx.x; b < d.buckets; b += blockDim.x) counts[b] = PLACE ? mine[b] + starts[b].x : 0u;
    __syncthreads();
    const uint32_t i0 = j << d.chunk_shift, i1 = min(d.n, i0 + (1u << d.chunk_shift));
    uint32_t kconst[9];
    recode_constant(d, kconst);
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        uint32_t t[9] = {0};
        if (!PLANES) {
            const uint4* sp = reinterpret_cast<const uint4*>(scalars + size_t(d.scalar_base + i) * 8);
            const uint4 a = sp[0], b = sp[1];
            const uint32_t s[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            recode_add(s, kconst, t);
        }
        for (uint32_t w = w0; w < w1; ++w) {
            uint32_t v, sign;
            if (PLANES) {
                const uint32_t e = planes[size_t(d.entry_base) + size_t(w) * d.n + i];
                v = e & 0x7fffu;
                sign = e >> 15;
            } else {
                const int c = int(window_width(d, w));
                const int32_t digit = int32_t(window_of(t, window_offset(d, w), c)) - int32_t((1u << (c - 1)) - 1);
                v = uint32_t(digit < 0 ? -digit : digit);
                sign = digit < 0 ? 1u : 0u;
            }
            if (v == 0) continue;
            const uint32_t pos = atomicAdd(&counts[v - 1], 1u);
            if (PLACE) {
                uint32_t pidx = point_index ? point_index[d.point_base + i] : d.point_base + i;
                if (d.table) {
                    const uint64_t p = uint64_t(pidx) * d.digit_windows + w;  // row, window -> table point
                    pidx = p < point_limit ? uint32_t(p) : point_limit;
                }
                if (pidx >= point_limit) {  // flagged for the host; point 0 keeps the access in bounds
                    flags[0] = 1u;
                    pidx = 0;
                }
                sorted[pos] = pidx | (sign << 31);
            }
        }
    }
    if (!PLACE) {
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < d.buckets; b += blockDim.x) mine[b] = counts[b];
    }
}

// Placing pass in two steps whose stores stay within short runs (group
// mode).  Placing every entry straight at its bucket slot, as above, leaves
// a block's entries spread over its window's whole entry range (about one
// entry per 128-byte line per block), so each line is written in pieces by
// many blocks at different times.  Here buckets are taken in groups of
// GROUP_BUCKETS consecutive buckets, whose entries are contiguous in the
// sorted order:
//   k_digits_group: the blocks of the counting pass again; the block's run
//     in a group starts at the group's first entry plus the block's prefixes
//     summed over the group's buckets, and every point goes to its group's
//     run (LDS atomic) as (MSM-local point index | bucket mod GROUP_BUCKETS
//     << 24 | sign << 31) in a staging array;
//   k_group_place: one block per group moves the group's staged entries
//     (one contiguous range) to their buckets' slots within the same range.
// Requires !table, buckets a multiple of GROUP_BUCKETS and n <= 2^24.
constexpr uint32_t GROUP_SHIFT = 6;
constexpr uint32_t GROUP_BUCKETS = 1u << GROUP_SHIFT;

// Entries a staging block sorts by group in LDS before writing them out, so
// that each group's run leaves as contiguous stores (16-byte aligned pieces
// of whole lines) instead of one scattered 4-byte store per entry.
constexpr uint32_t GROUP_STAGE = 15360;

static __global__ void __launch_bounds__(MSM_DIGIT_THREADS)
    k_digits_group(const MsmDesc* __restrict__ descs, uint32_t m, uint32_t grid, const uint16_t* __restrict__ planes,
                   const uint32_t* __restrict__ hist, const uint2* __restrict__ offsets,
                   uint32_t* __restrict__ staged) {
    constexpr uint32_t GROUPS = MSM_LDS_BUCKETS / GROUP_BUCKETS;
    __shared__ uint32_t cursor[GROUPS];  // this block's next slot in each group's run
    __shared__ uint32_t lcount[GROUPS], lfill[GROUPS];
    __shared__ uint32_t ents[GROUP_STAGE];
    const uint32_t g = xcd_grouped_block(blockIdx.x, grid);
    const MsmDesc d = descs[find_desc<&MsmDesc::hist_block_base>(descs, m, g)];
    const uint32_t local = g - d.hist_block_base;
    const uint32_t j = local / d.digit_windows, w = local % d.digit_windows;
    const uint32_t* mine = hist + d.hist_base + size_t(local) * d.buckets;
    const uint2* starts = offsets + d.bucket_base + w * d.buckets;
    const uint32_t groups = d.buckets >> GROUP_SHIFT, lane = threadIdx.x & 63, wave = threadIdx.x >> 6,
                   waves = blockDim.x >> 6;
    for (uint32_t s = wave; s < groups; s += waves) {  // one wave per group
        const uint32_t sum = wave_sum(mine[s * GROUP_BUCKETS + lane]);
        if (lane == 0) cursor[s] = starts[s * GROUP_BUCKETS].x + sum;
    }
    const uint32_t i0 = j << d.chunk_shift, i1 = min(d.n, i0 + (1u << d.chunk_shift));
    const uint16_t* plane = planes + size_t(d.entry_base) + size_t(w) * d.n;
    for (uint32_t r0 = i0; r0 < i1; r0 += GROUP_STAGE) {
        const uint32_t r1 = min(i1, r0 + GROUP_STAGE);
        for (uint32_t s = threadIdx.x; s < groups; s += blockDim.x) lcount[s] = 0;
        __syncthreads();
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            const uint32_t v = plane[i] & 0x7fffu;
            if (v) atomicAdd(&lcount[(v - 1u) >> GROUP_SHIFT], 1u);
        }
        __syncthreads();
        if (wave == 0) {  // exclusive prefix of the group counts (up to four groups a lane)
            uint32_t c[4], sum = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t s = lane * 4 + q;
                c[q] = s < groups ? lcount[s] : 0u;
                sum += c[q];
            }
            uint32_t incl = sum;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(incl, off, 64);
                if (lane >= uint32_t(off)) incl += t;
            }
            uint32_t at = incl - sum;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t s = lane * 4 + q;
                if (s < groups) lfill[s] = at;
                at += c[q];
            }
        }
        __syncthreads();
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            const uint32_t e = plane[i];
            if ((e & 0x7fffu) == 0) continue;
            const uint32_t bucket = (e & 0x7fffu) - 1u;
            const uint32_t pos = atomicAdd(&lfill[bucket >> GROUP_SHIFT], 1u);
            ents[pos] = i | ((bucket & (GROUP_BUCKETS - 1u)) << 24) | ((e >> 15) << 31);
        }
        __syncthreads();
        for (uint32_t s = wave; s < groups; s += waves) {  // one wave copies a group's run
            const uint32_t n = lcount[s], src = lfill[s] - n, dst = cursor[s];
            for (uint32_t k = lane; k < n; k += 64) staged[dst + k] = ents[src + k];
            if (lane == 0) cursor[s] = dst + n;
        }
        __syncthreads();
    }
}

// Entries k_group_place sorts by bucket in LDS per round, so that each
// bucket's share of the round leaves as one contiguous piece.
constexpr uint32_t PLACE_STAGE = 4096;

static __global__ void __launch_bounds__(256)
    k_group_place(const MsmDesc* __restrict__ descs, uint32_t m, const uint32_t* __restrict__ counts,
                  const uint2* __restrict__ offsets, const uint32_t* __restrict__ staged,
                  const uint32_t* __restrict__ point_index, uint32_t point_limit, uint32_t* __restrict__ flags,
                  uint32_t* __restrict__ sorted) {
    __shared__ uint32_t cursor[GROUP_BUCKETS], lcount[GROUP_BUCKETS], lfill[GROUP_BUCKETS];
    __shared__ uint32_t ents[PLACE_STAGE];
    __shared__ uint32_t range[2];
    const uint32_t gb0 = blockIdx.x * GROUP_BUCKETS;
    const MsmDesc d = descs[find_desc<&MsmDesc::bucket_base>(descs, m, gb0)];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    if (wave == 0) {  // one bucket per lane
        const uint32_t start = offsets[gb0 + lane].x;
        cursor[lane] = start;
        const uint32_t len = wave_sum(counts[gb0 + lane]);
        if (lane == 0) {
            range[0] = start;
            range[1] = len;
        }
    }
    __syncthreads();
    const uint32_t start = range[0], len = range[1];
    for (uint32_t r0 = 0; r0 < len; r0 += PLACE_STAGE) {
        const uint32_t r1 = min(len, r0 + PLACE_STAGE);
        if (threadIdx.x < GROUP_BUCKETS) lcount[threadIdx.x] = 0;
        __syncthreads();
        for (uint32_t k = r0 + threadIdx.x; k < r1; k += blockDim.x)
            atomicAdd(&lcount[(staged[start + k] >> 24) & (GROUP_BUCKETS - 1u)], 1u);
        __syncthreads();
        if (wave == 0) {  // exclusive prefix of the bucket counts, one a lane
            const uint32_t c = lcount[lane];
            uint32_t incl = c;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(incl, off, 64);
                if (lane >= uint32_t(off)) incl += t;
            }
            lfill[lane] = incl - c;
        }
        __syncthreads();
        for (uint32_t k = r0 + threadIdx.x; k < r1; k += blockDim.x) {
            const uint32_t e = staged[start + k];
            const uint32_t i = e & 0xffffffu;
            uint32_t pidx = point_index ? point_index[d.point_base + i] : d.point_base + i;
            if (pidx >= point_limit) {  // flagged for the host; point 0 keeps the access in bounds
                flags[0] = 1u;
                pidx = 0;
            }
            const uint32_t pos = atomicAdd(&lfill[(e >> 24) & (GROUP_BUCKETS - 1u)], 1u);
            ents[pos] = pidx | (e & 0x80000000u);
        }
        __syncthreads();
        for (uint32_t b = wave; b < GROUP_BUCKETS; b += waves) {  // one wave copies a bucket's piece
            const uint32_t n = lcount[b], src = lfill[b] - n, dst = cursor[b];
            for (uint32_t k = lane; k < n; k += 64) sorted[dst + k] = ents[src + k];
            if (lane == 0) cursor[b] = dst + n;
        }
        __syncthreads();
    }
}

// One thread per bucket: exclusive prefix of the bucket's per-block counts
// over the blocks feeding it (in place), the bucket's total, and the task
// classes of the block's buckets added to the global class counts.
static __global__ void __launch_bounds__(256) k_hist_prefix(const MsmDesc* __restrict__ descs, uint32_t m,
                                                           uint32_t total_buckets, uint32_t* __restrict__ hist,
                                                           uint32_t* __restrict__ counts, uint32_t tp,
                                                           uint32_t* __restrict__ total) {
    __shared__ uint32_t h[TASK_CLASS_MAX], hs[HEAVY_TASKS];
    for (uint32_t i = threadIdx.x; i < tp; i += blockDim.x) h[i] = 0;
    if (threadIdx.x < HEAVY_TASKS) hs[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t run = 0;
    if (gb < total_buckets) {
        const MsmDesc d = descs[find_desc<&MsmDesc::bucket_base>(descs, m, gb)];
        const uint32_t local = gb - d.bucket_base, row = local / d.buckets, b = local % d.buckets;
        // Normal mode: bucket row w is fed by blocks (chunk j, window w); table
        // mode: the single row is fed by every block of the MSM.
        const uint32_t nblk = d.chunks;
        for (uint32_t k = 0; k < nblk; ++k) {
            const uint32_t blk = d.table ? k : k * d.digit_windows + row;
            uint32_t* p = hist + d.hist_base + size_t(blk) * d.buckets + b;
            const uint32_t v = *p;
            *p = run;
            run += v;
        }
        counts[gb] = run;
    }
    task_class_add(run, tp, h, hs);
    __syncthreads();
    task_class_flush(tp, h, hs, total);
}

static __global__ void k_scatter(const MsmDesc* __restrict__ descs, uint32_t m, uint32_t total_entries,
                                 const uint32_t* __restrict__ digits, const uint32_t* __restrict__ ranks,
                                 const uint2* __restrict__ offsets, const uint32_t* __restrict__ point_index,
                                 uint32_t point_limit, uint32_t* __restrict__ flags, uint32_t* __restrict__ sorted) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total_entries) return;
    const uint32_t dig = digits[e];
    if (dig == NO_DIGIT) return;
    const MsmDesc d = descs[find_desc<&MsmDesc::entry_base>(descs, m, e)];
    const uint32_t local = e - d.entry_base, w = local / d.n, i = local % d.n;
    const uint32_t pos = ranks[e] + offsets[d.bucket_base + (d.table ? 0u : w) * d.buckets + (dig >> 1)].x;
    uint32_t pidx = point_index ? point_index[d.point_base + i] : d.point_base + i;
    if (d.table) {
        const uint64_t p = uint64_t(pidx) * d.digit_windows + w;  // row, window -> table point
        pidx = p < point_limit ? uint32_t(p) : point_limit;
    }
    if (pidx >= point_limit) {  // flagged for the host; point 0 keeps the access in bounds
        flags[0] = 1u;
        pidx = 0;
    }
    sorted[pos] = pidx | ((dig & 1u) << 31);
}

// Exclusive scan of (count, tasks(count)) pairs in three phases.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint2 pair_of(uint32_t count, uint32_t task_points) {
    return make_uint2(count, (count + task_points - 1) / task_points);
}

__device__ __forceinline__ uint2 block_exclusive_scan(uint2 v, uint2* smem, uint2& total) {
    const int t = threadIdx.x;
    smem[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint2 add = make_uint2(0, 0);
        if (t >= off) add = smem[t - off];
        __syncthreads();
        smem[t].x += add.x;
        smem[t].y += add.y;
        __syncthreads();
    }
    const uint2 incl = smem[t];
    total = smem[SCAN_THREADS - 1];
    __syncthreads();
    return make_uint2(incl.x - v.x, incl.y - v.y);
}

static __global__ void k_scan_tiles(const uint32_t* __restrict__ counts, uint32_t m, uint32_t task_points,
                             uint2* __restrict__ tile_sums) {
    __shared__ uint2 smem[SCAN_THREADS];
    const size_t base = size_t(blockIdx.x) * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint2 local = make_uint2(0, 0);
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < m) {
            const uint2 p = pair_of(counts[base + k], task_points);
            local.x += p.x;
            local.y += p.y;
        }
    }
    uint2 total;
    block_exclusive_scan(local, smem, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place; writes the grand
// total (points, tasks) to total[0..1] and the task-class cursors.
static __global__ void k_scan_sums(uint2* __restrict__ tile_sums, uint32_t tiles, uint32_t tp,
                                   uint32_t* __restrict__ total) {
    __shared__ uint2 smem[SCAN_THREADS];
    uint2 carry = make_uint2(0, 0);
    for (uint32_t start = 0; start < tiles; start += SCAN_THREADS) {
        const uint32_t i = start + threadIdx.x;
        const uint2 v = i < tiles ? tile_sums[i] : make_uint2(0, 0);
        uint2 blk;
        const uint2 ex = block_exclusive_scan(v, smem, blk);
        if (i < tiles) tile_sums[i] = make_uint2(ex.x + carry.x, ex.y + carry.y);
        carry.x += blk.x;
        carry.y += blk.y;
    }
    if (threadIdx.x == 0) {
        total[0] = carry.x;
        total[1] = carry.y;
        uint32_t listed = 0;  // bucket-sum class cursors
        for (uint32_t c = 0; c < HEAVY_TASKS; ++c) {
            total[TOTAL_SUM_CURSORS + c] = listed;
            listed += total[TOTAL_SUM_COUNTS + c];
        }
        total[TOTAL_SUM_LISTED] = listed;
    }
    // Task-class cursors: exclusive prefix of the class counts (classes
    // [4t, 4t + 4) per thread).
    constexpr uint32_t PER = TASK_CLASS_MAX / SCAN_THREADS;
    uint32_t v[PER], sum = 0;
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t c = threadIdx.x * PER + k;
        v[k] = c < tp ? total[TOTAL_CLASS_COUNTS + c] : 0u;
        sum += v[k];
    }
    uint2 all;
    uint32_t run = block_exclusive_scan(make_uint2(sum, 0u), smem, all).x;
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t c = threadIdx.x * PER + k;
        if (c < tp) total[TOTAL_CLASS_CURSORS + c] = run;
        run += v[k];
    }
}

static __global__ void k_scan_apply(const uint32_t* __restrict__ counts, uint32_t m, uint32_t task_points,
                             const uint2* __restrict__ tile_sums, uint2* __restrict__ offsets) {
    __shared__ uint2 smem[SCAN_THREADS];
    const size_t base = size_t(blockIdx.x) * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint2 vals[SCAN_ITEMS];
    uint2 local = make_uint2(0, 0);
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        vals[k] = base + k < m ? pair_of(counts[base + k], task_points) : make_uint2(0, 0);
        local.x += vals[k].x;
        local.y += vals[k].y;
    }
    uint2 total;
    uint2 run = block_exclusive_scan(local, smem, total);
    const uint2 tile = tile_sums[blockIdx.x];
    run.x += tile.x;
    run.y += tile.y;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < m) offsets[base + k] = run;
        run.x += vals[k].x;
        run.y += vals[k].y;
    }
}

// A task's partial sum keeps its bucket-order slot (offsets[b].y + k), so
// the bucket sums find a bucket's partials together.
// The task kernels give each thread `per` buckets (strided by the block
// size, so loads stay coalesced), about 256 blocks in all.
constexpr uint32_t TASK_THREADS = 256;

inline uint32_t task_buckets_per_thread(uint32_t buckets) {
    const uint32_t per = uint32_t((uint64_t(buckets) + 256 * TASK_THREADS - 1) / (256 * TASK_THREADS));
    return std::max<uint32_t>(1, std::min<uint32_t>(16, per));
}

// Class counts of `per` buckets per thread (the path without the LDS
// histogram pass, whose k_hist_prefix adds them itself).
static __global__ void __launch_bounds__(TASK_THREADS) k_task_classes(const uint32_t* __restrict__ counts, uint32_t m,
                                                                     uint32_t tp, uint32_t per,
                                                                     uint32_t* __restrict__ total) {
    __shared__ uint32_t h[TASK_CLASS_MAX], hs[HEAVY_TASKS];
    for (uint32_t i = threadIdx.x; i < tp; i += TASK_THREADS) h[i] = 0;
    if (threadIdx.x < HEAVY_TASKS) hs[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TASK_THREADS * per + threadIdx.x;
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t wb = base + k * TASK_THREADS;
        task_class_add(wb < m ? counts[wb] : 0u, tp, h, hs);
    }
    __syncthreads();
    task_class_flush(tp, h, hs, total);
}

// Every bucket's tasks {first entry, points, partial slot} at positions of
// their classes: the block takes one range per class, each thread one run
// of full tasks for all its buckets.
static __global__ void __launch_bounds__(TASK_THREADS) k_tasks(const uint32_t* __restrict__ counts,
                                                              const uint2* __restrict__ offsets, uint32_t m,
                                                              uint32_t tp, uint32_t per, uint4* __restrict__ tasks,
                                                              uint32_t* __restrict__ heavy,
                                                              uint32_t* __restrict__ order,
                                                              uint32_t* __restrict__ total) {
    __shared__ uint32_t h[TASK_CLASS_MAX], hs[HEAVY_TASKS];
    __shared__ uint32_t wave_full[TASK_THREADS / 64];
    for (uint32_t i = threadIdx.x; i < tp; i += TASK_THREADS) h[i] = 0;
    if (threadIdx.x < HEAVY_TASKS) hs[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TASK_THREADS * per + threadIdx.x;
    uint32_t full = 0;
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t wb = base + k * TASK_THREADS;
        const uint32_t cnt = wb < m ? counts[wb] : 0u, r = cnt % tp, sc = sum_class(cnt, tp);
        full += cnt / tp;
        if (r) atomicAdd(&h[tp - r], 1u);
        if (sc < HEAVY_TASKS) atomicAdd(&hs[sc], 1u);
    }
    // this thread's first full task: block range + waves before + lanes before
    uint32_t incl = full;
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wave_full[threadIdx.x >> 6] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t all = 0;
        for (uint32_t w = 0; w < TASK_THREADS / 64; ++w) {
            const uint32_t c = wave_full[w];
            wave_full[w] = all;
            all += c;
        }
        h[0] = all;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tp; i += TASK_THREADS)
        if (h[i]) h[i] = atomicAdd(&total[TOTAL_CLASS_CURSORS + i], h[i]);
    if (threadIdx.x < HEAVY_TASKS && hs[threadIdx.x])
        hs[threadIdx.x] = atomicAdd(&total[TOTAL_SUM_CURSORS + threadIdx.x], hs[threadIdx.x]);
    __syncthreads();
    uint32_t pos = h[0] + wave_full[threadIdx.x >> 6] + incl - full;
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t wb = base + k * TASK_THREADS;
        if (wb >= m) break;
        const uint32_t cnt = counts[wb], q = cnt / tp, r = cnt % tp;
        const uint2 off = offsets[wb];
        for (uint32_t j = 0; j < q; ++j) tasks[pos++] = make_uint4(off.x + j * tp, tp, off.y + j, 0u);
        if (r) tasks[atomicAdd(&h[tp - r], 1u)] = make_uint4(off.x + q * tp, r, off.y + q, 0u);
        if (q + (r ? 1u : 0u) > HEAVY_TASKS) heavy[atomicAdd(&total[2], 1u)] = wb;
        const uint32_t sc = sum_class(cnt, tp);
        if (sc < HEAVY_TASKS) order[atomicAdd(&hs[sc], 1u)] = wb;
    }
}

template <class F>
__global__ void __launch_bounds__(256) k_accumulate(const uint32_t* __restrict__ points,
                                                    const uint32_t* __restrict__ sorted,
                                                    const uint4* __restrict__ tasks,
                                                    const uint32_t* __restrict__ total,
                                                    uint32_t* __restrict__ partials) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total[1]) return;
    const uint4 task = tasks[t];  // first entry, points, partial slot
    constexpr int AW = PointLayout<F>::AFFINE_WORDS;
    Xyzz<F> acc = xyzz_inf<F>();
    // Software pipeline: the next point's gather is in flight while the
    // current point is added.
    uint32_t entry = task.y ? sorted[task.x] : 0u;
    uint4 raw[AW / 4];
    if (task.y) {
        const uint4* q = reinterpret_cast<const uint4*>(points + size_t(entry & 0x7fffffffu) * AW);
#pragma unroll
        for (int k = 0; k < AW / 4; ++k) raw[k] = q[k];
    }
    for (uint32_t e = 0; e < task.y; ++e) {
        uint32_t cur[AW];
#pragma unroll
        for (int k = 0; k < AW / 4; ++k) {
            cur[4 * k] = raw[k].x;
            cur[4 * k + 1] = raw[k].y;
            cur[4 * k + 2] = raw[k].z;
            cur[4 * k + 3] = raw[k].w;
        }
        const uint32_t sign = entry >> 31;
        if (e + 1 < task.y) {
            entry = sorted[task.x + e + 1];
            const uint4* q = reinterpret_cast<const uint4*>(points + size_t(entry & 0x7fffffffu) * AW);
#pragma unroll
            for (int k = 0; k < AW / 4; ++k) raw[k] = q[k];
        }
        const Affine<F> p = Mem<F>::load_affine(cur);
        if (p.inf) continue;
        acc = xyzz_madd_signed<F>(acc, p.x, p.y, sign != 0);
    }
    Mem<F>::store_xyzz(partials + size_t(task.z) * PointLayout<F>::XYZZ_WORDS, acc);
}

// One block per heavy bucket (grid-strided): strided serial sums, then an
// LDS tree; the bucket total overwrites the bucket's first task partial.
template <class F>
__global__ void __launch_bounds__(256) k_heavy(const uint32_t* __restrict__ heavy,
                                               const uint32_t* __restrict__ heavy_count,
                                               const uint32_t* __restrict__ counts,
                                               const uint2* __restrict__ offsets, uint32_t task_points,
                                               uint32_t* __restrict__ partials) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    __shared__ uint4 lds[256 * XW / 4];
    uint32_t* slot = reinterpret_cast<uint32_t*>(lds) + threadIdx.x * XW;
    const uint32_t nheavy = *heavy_count;
    for (uint32_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const uint32_t wb = heavy[h];
        const uint32_t t0 = offsets[wb].y;
        const uint32_t nt = (counts[wb] + task_points - 1) / task_points;
        Xyzz<F> acc = xyzz_inf<F>();
        for (uint32_t k = threadIdx.x; k < nt; k += 256)
            acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(partials + size_t(t0 + k) * XW));
        Mem<F>::store_xyzz(slot, acc);
        __syncthreads();
        for (uint32_t s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(slot + s * XW));
                Mem<F>::store_xyzz(slot, acc);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) Mem<F>::store_xyzz(partials + size_t(t0) * XW, acc);
        __syncthreads();
    }
}


// One thread per bucket of 2..HEAVY_TASKS tasks (listed by k_tasks in order
// of decreasing task count): the bucket's task partials summed into its
// first task's partial (heavier buckets are summed by k_heavy), so the
// chunk reduction reads one value per bucket.
template <class F>
__global__ void __launch_bounds__(256) k_bucket_sums(const uint32_t* __restrict__ counts,
                                                     const uint2* __restrict__ offsets,
                                                     const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ total, uint32_t task_points,
                                                     uint32_t* __restrict__ partials) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total[TOTAL_SUM_LISTED]) return;
    const uint32_t wb = order[i];
    const uint32_t nt = (counts[wb] + task_points - 1) / task_points;
    const uint32_t t0 = offsets[wb].y;
    Xyzz<F> s = Mem<F>::load_xyzz(partials + size_t(t0) * XW);
    for (uint32_t k = 1; k < nt; ++k) s = xyzz_add<F>(s, Mem<F>::load_xyzz(partials + size_t(t0 + k) * XW));
    Mem<F>::store_xyzz(partials + size_t(t0) * XW, s);
}

// One thread per chunk of L buckets of one window:
//   out = sum_i (i + 1) S_{jL+i} + (jL) * sum_i S_{jL+i}
// where S_b is the sum of bucket b's task partials.
// SPLIT (chunk-plane mode): write U = sum_i (i + 1) S_{jL+i} and
// T = sum_i S_{jL+i} as out[2g], out[2g + 1] instead, leaving the offset
// term jL * T to the chunk planes.
template <class F, bool SPLIT = false>
__global__ void __launch_bounds__(256) k_reduce_chunks(const MsmDesc* __restrict__ descs, uint32_t m,
                                                       uint32_t total_chunks, const uint32_t* __restrict__ counts,
                                                       const uint2* __restrict__ offsets,
                                                       const uint32_t* __restrict__ partials,
                                                       uint32_t* __restrict__ out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total_chunks) return;
    const MsmDesc d = descs[find_desc<&MsmDesc::chunk_base>(descs, m, g)];
    const int chunk = int(d.chunk);
    const uint32_t per_window = d.buckets / uint32_t(chunk);
    const uint32_t local = g - d.chunk_base, w = local / per_window, j = local % per_window;
    Xyzz<F> run = xyzz_inf<F>(), acc = xyzz_inf<F>();
    for (int i = chunk - 1; i >= 0; --i) {
        const size_t wb = size_t(d.bucket_base) + size_t(w) * d.buckets + size_t(j) * chunk + i;
        const uint32_t cnt = counts[wb];
        if (cnt)  // the bucket total sits in its first task's partial (k_bucket_sums, k_heavy)
            run = xyzz_add<F>(run, Mem<F>::load_xyzz(partials + size_t(offsets[wb].y) * PointLayout<F>::XYZZ_WORDS));
        acc = xyzz_add<F>(acc, run);
    }
    if constexpr (SPLIT) {
        Mem<F>::store_xyzz(out + size_t(2 * g) * PointLayout<F>::XYZZ_WORDS, acc);
        Mem<F>::store_xyzz(out + size_t(2 * g + 1) * PointLayout<F>::XYZZ_WORDS, run);
        return;
    }
    const uint32_t offset = j * uint32_t(chunk);
    if (offset) {
        const uint32_t k[1] = {offset};
        acc = xyzz_add<F>(acc, xyzz_mul_words<F, 1>(run, k, 32 - __clz(offset)));
    }
    Mem<F>::store_xyzz(out + size_t(g) * PointLayout<F>::XYZZ_WORDS, acc);
}

// One block per window of the batch: strided sums of the window's chunk
// values, then an LDS tree; thread 0 writes the window value.
template <class F>
struct TreeThreads {
    // 64 KiB of LDS partials: 512 G1 or 256 G2 XYZZ points per block.
    static constexpr int value = 65536 / (PointLayout<F>::XYZZ_WORDS * 4);
};

template <class F, int T>
__global__ void __launch_bounds__(T)
    k_window_tree(const MsmDesc* __restrict__ descs, uint32_t m, const uint32_t* __restrict__ chunk_vals,
                  uint32_t* __restrict__ out) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    __shared__ uint4 lds[T * XW / 4];
    uint32_t* slot = reinterpret_cast<uint32_t*>(lds) + threadIdx.x * XW;
    const uint32_t gw = blockIdx.x;
    const MsmDesc d = descs[find_desc<&MsmDesc::window_base>(descs, m, gw)];
    const uint32_t w = gw - d.window_base, per_window = d.buckets / d.chunk;
    const size_t base = size_t(d.chunk_base) + size_t(w) * per_window;
    Xyzz<F> acc = xyzz_inf<F>();
    for (uint32_t k = threadIdx.x; k < per_window; k += T)
        acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(chunk_vals + (base + k) * XW));
    Mem<F>::store_xyzz(slot, acc);
    __syncthreads();
    for (uint32_t s = T / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(slot + s * XW));
            Mem<F>::store_xyzz(slot, acc);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) Mem<F>::store_xyzz(out + size_t(gw) * XW, acc);
}

// ---- plane mode: bit-plane sums of the bucket weights ----------------------
//
// For one window, sum_i (i + 1) S_i = sum_k 2^k P_k with P_k the sum of the
// bucket sums S_i whose weight v = i + 1 has bit k set.  Each P_k is a plain
// parallel sum (depth ~log of the buckets), where the running-sum form is a
// chain of two adds per bucket plus a scalar multiplication; the host forms
// sum_k 2^k P_k.  The weights with bit k set are enumerated directly:
// v_k(j) = ((j >> k) << (k + 1)) | 2^k | (j & (2^k - 1)), increasing in j.
constexpr int PLANE_THREADS = 256;

__host__ __device__ __forceinline__ uint32_t plane_bits(uint32_t buckets) {
    uint32_t b = 0;
    while (b < 32 && (buckets >> b)) ++b;
    return b;  // bit length of the largest weight
}

// Weights in [1, buckets] with bit k set.
__host__ __device__ __forceinline__ uint32_t plane_count(uint32_t buckets, uint32_t k) {
    const uint64_t period = uint64_t(2) << k, half = uint64_t(1) << k, top = uint64_t(buckets) + 1;
    const uint64_t rest = top % period;
    return uint32_t(top / period * half + (rest > half ? rest - half : 0));
}

// Block g: slice g % slices of plane g / slices (planes numbered per MSM
// from plane_base, window-major).  out: one XYZZ partial per block.
template <class F>
__global__ void __launch_bounds__(PLANE_THREADS)
    k_plane_sums(const MsmDesc* __restrict__ descs, uint32_t m, uint32_t slices, const uint32_t* __restrict__ counts,
                 const uint2* __restrict__ offsets, const uint32_t* __restrict__ partials, uint32_t* __restrict__ out) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    __shared__ uint4 lds[PLANE_THREADS * XW / 4];
    uint32_t* slot = reinterpret_cast<uint32_t*>(lds) + threadIdx.x * XW;
    const uint32_t g = blockIdx.x, plane = g / slices, s = g % slices;
    const MsmDesc d = descs[find_desc<&MsmDesc::plane_base>(descs, m, plane)];
    const uint32_t pb = plane_bits(d.buckets);
    const uint32_t local = plane - d.plane_base, w = local / pb, k = local % pb;
    const uint32_t count = plane_count(d.buckets, k);
    const uint32_t lo = uint32_t(uint64_t(count) * s / slices), hi = uint32_t(uint64_t(count) * (s + 1) / slices);
    const size_t wbase = size_t(d.bucket_base) + size_t(w) * d.buckets;
    Xyzz<F> acc = xyzz_inf<F>();
    for (uint32_t j = lo + threadIdx.x; j < hi; j += PLANE_THREADS) {
        const uint32_t v = ((j >> k) << (k + 1)) | (1u << k) | (j & ((1u << k) - 1u));
        const size_t wb = wbase + (v - 1);
        if (counts[wb])  // the bucket total sits in its first task's partial (k_bucket_sums, k_heavy)
            acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(partials + size_t(offsets[wb].y) * XW));
    }
    Mem<F>::store_xyzz(slot, acc);
    __syncthreads();
    for (uint32_t h = PLANE_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(slot + h * XW));
            Mem<F>::store_xyzz(slot, acc);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) Mem<F>::store_xyzz(out + size_t(g) * XW, acc);
}

// One thread per plane: the sum of its slices.
template <class F>
__global__ void __launch_bounds__(64) k_plane_final(const uint32_t* __restrict__ parts, uint32_t planes,
                                                    uint32_t slices, uint32_t* __restrict__ out) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= planes) return;
    Xyzz<F> acc = Mem<F>::load_xyzz(parts + size_t(p) * slices * XW);
    for (uint32_t s = 1; s < slices; ++s) acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(parts + (size_t(p) * slices + s) * XW));
    Mem<F>::store_xyzz(out + size_t(p) * XW, acc);
}

// Plane mode does about popcount-many adds per bucket where the running sums
// do two, so it pays only while its launch fits one wave round of the chip
// (latency-bound chains); larger bucket sets keep the running-sum reduction.
constexpr uint32_t PLANE_MAX_THREADS = 131072;  // 256 CUs x 4 SIMDs x 2 waves x 64 lanes

// Switch a plan to plane mode (before msm_batch_prepare) when that pays:
// plane bases, and slices so that a thread adds at most about 8 buckets
// before the LDS tree.  Returns whether the plan is in plane mode.
inline bool msm_plan_planes(MsmBatchPlan& plan) {
    uint32_t total = 0, widest = 1;
    for (const MsmDesc& d : plan.descs) {
        total += d.windows * plane_bits(d.buckets);
        widest = std::max(widest, plane_count(d.buckets, 0));
    }
    const uint32_t slices = std::min<uint32_t>(16, std::max<uint32_t>(1, widest / (PLANE_THREADS * 8)));
    if (uint64_t(total) * slices * PLANE_THREADS > PLANE_MAX_THREADS) return false;
    total = 0;
    for (MsmDesc& d : plan.descs) {
        d.plane_base = total;
        total += d.windows * plane_bits(d.buckets);
    }
    plan.planes = true;
    plan.total_planes = total;
    plan.plane_slices = slices;
    return true;
}

// ---- chunk-plane mode: bit planes over the chunk index --------------------
//
// For larger bucket sets: the chunk running sums (k_reduce_chunks<SPLIT>)
// give U_j and T_j per chunk j of L buckets, and the window value is
// sum_j U_j + L * sum_k 2^k Q_k with Q_k the sum of the T_j whose index j has
// bit k set, k < log2(J), J = buckets / L chunks.  Each sum is cut into
// slices of at most CHUNK_PLANE_SPAN values per thread; per window the
// outputs are [U slices][Q_0 slices][Q_1 slices]..., summed on the host.
constexpr uint32_t CHUNK_PLANE_SPAN = 4;

struct ChunkPlaneShape {
    uint32_t J, qb, su, sq, per_window;  // chunks, Q planes, U / Q slices, outputs per window
};

__host__ __device__ __forceinline__ ChunkPlaneShape chunk_plane_shape(uint32_t buckets, uint32_t chunk) {
    ChunkPlaneShape c;
    c.J = buckets / chunk;
    c.qb = plane_bits(c.J) - 1;  // log2(J), J a power of two
    const uint32_t span = PLANE_THREADS * CHUNK_PLANE_SPAN;
    c.su = (c.J + span - 1) / span;
    c.sq = (c.J / 2 + span - 1) / span;
    if (c.sq == 0) c.sq = 1;
    c.per_window = c.su + c.qb * c.sq;
    return c;
}

// Block g: one slice of one plane (plane_base numbers this MSM's outputs).
template <class F>
__global__ void __launch_bounds__(PLANE_THREADS)
    k_chunk_planes(const MsmDesc* __restrict__ descs, uint32_t m, const uint32_t* __restrict__ chunk_vals,
                   uint32_t* __restrict__ out) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    __shared__ uint4 lds[PLANE_THREADS * XW / 4];
    uint32_t* slot = reinterpret_cast<uint32_t*>(lds) + threadIdx.x * XW;
    const uint32_t g = blockIdx.x;
    const MsmDesc d = descs[find_desc<&MsmDesc::plane_base>(descs, m, g)];
    const ChunkPlaneShape sh = chunk_plane_shape(d.buckets, d.chunk);
    const uint32_t local = g - d.plane_base, w = local / sh.per_window, r = local % sh.per_window;
    const size_t base = size_t(d.chunk_base) + size_t(w) * sh.J;  // chunk index of (window w, j = 0)
    Xyzz<F> acc = xyzz_inf<F>();
    if (r < sh.su) {  // a slice of sum_j U_j
        const uint32_t lo = uint32_t(uint64_t(sh.J) * r / sh.su), hi = uint32_t(uint64_t(sh.J) * (r + 1) / sh.su);
        for (uint32_t j = lo + threadIdx.x; j < hi; j += PLANE_THREADS)
            acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(chunk_vals + (2 * (base + j)) * XW));
    } else {  // a slice of Q_k
        const uint32_t k = (r - sh.su) / sh.sq, s = (r - sh.su) % sh.sq, cnt = sh.J / 2;
        const uint32_t lo = uint32_t(uint64_t(cnt) * s / sh.sq), hi = uint32_t(uint64_t(cnt) * (s + 1) / sh.sq);
        for (uint32_t t = lo + threadIdx.x; t < hi; t += PLANE_THREADS) {
            const uint32_t j = ((t >> k) << (k + 1)) | (1u << k) | (t & ((1u << k) - 1u));
            acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(chunk_vals + (2 * (base + j) + 1) * XW));
        }
    }
    Mem<F>::store_xyzz(slot, acc);
    __syncthreads();
    for (uint32_t h = PLANE_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(slot + h * XW));
            Mem<F>::store_xyzz(slot, acc);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) Mem<F>::store_xyzz(out + size_t(g) * XW, acc);
}

// Switch a plan to chunk-plane mode (before msm_batch_prepare); every
// descriptor's chunk must divide its bucket count (powers of two).
inline void msm_plan_chunk_planes(MsmBatchPlan& plan) {
    uint32_t total = 0;
    for (MsmDesc& d : plan.descs) {
        require(d.chunk && d.buckets % d.chunk == 0 && ((d.buckets / d.chunk) & (d.buckets / d.chunk - 1)) == 0,
                "chunk planes need power-of-two chunks per window");
        d.plane_base = total;
        total += d.windows * chunk_plane_shape(d.buckets, d.chunk).per_window;
    }
    plan.chunk_planes = true;
    plan.total_planes = total;
    plan.plane_slices = 1;
}

// Horner over one MSM's window values and conversion to affine, one lane per
// MSM (the batch's many small MSMs finish in parallel on the device).
template <class F>
__global__ void __launch_bounds__(64) k_finalize(const MsmDesc* __restrict__ descs, uint32_t m,
                                                 const uint32_t* __restrict__ windows, uint8_t* __restrict__ out,
                                                 uint32_t* __restrict__ slot_points, const uint32_t* __restrict__ gate) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    constexpr int AW = PointLayout<F>::AFFINE_WORDS;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const MsmDesc d = descs[i];
    const uint32_t* base = windows + size_t(d.window_base) * XW;
    Xyzz<F> acc = Mem<F>::load_xyzz(base + size_t(d.windows - 1) * XW);
    for (int w = int(d.windows) - 2; w >= 0; --w) {
        for (uint32_t k = 0; k < window_width(d, uint32_t(w)); ++k) acc = xyzz_dbl<F>(acc);
        acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(base + size_t(w) * XW));
    }
    const Affine<F> a = xyzz_to_affine<F>(acc);
    Fq comp[AW / 8];
    if constexpr (AW == 16) {
        comp[0] = a.x;
        comp[1] = a.y;
    } else {
        comp[0] = a.x.c0;
        comp[1] = a.x.c1;
        comp[2] = a.y.c0;
        comp[3] = a.y.c1;
    }
    uint32_t* o = reinterpret_cast<uint32_t*>(out) + size_t(d.out_index) * AW;
    // A flagged batch (an indexed point out of range) leaves every slot as it was.
    const bool store = slot_points && d.result_slot != MSM_NO_SLOT && gate[0] == 0;
    uint32_t* sp = store ? slot_points + size_t(d.result_slot) * AW : nullptr;
    for (int f = 0; f < AW / 8; ++f) {
        uint32_t cw[8] = {0}, mw[8] = {0};
        if (!a.inf) {
            from_mont(comp[f], cw);
            pack(canon(comp[f]), mw);
        }
        for (int k = 0; k < 8; ++k) o[f * 8 + k] = cw[k];
        if (sp)
            for (int k = 0; k < 8; ++k) sp[f * 8 + k] = mw[k];
    }
}

// Canonical little-endian bytes -> Montgomery words; all-zero stays zero.
template <class F>
__global__ void k_to_mont(const uint8_t* __restrict__ canonical, uint32_t* __restrict__ mont, uint32_t n) {
    constexpr int W = PointLayout<F>::AFFINE_WORDS;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(canonical) + size_t(i) * W;
    uint32_t* dst = mont + size_t(i) * W;
    uint32_t any = 0;
    for (int k = 0; k < W; ++k) any |= src[k];
    for (int f = 0; f < W / 8; ++f) {
        uint32_t w[8], o[8];
        for (int k = 0; k < 8; ++k) w[k] = src[f * 8 + k];
        if (any) pack(canon(to_mont<FqParams>(w)), o);
        else for (int k = 0; k < 8; ++k) o[k] = 0;
        for (int k = 0; k < 8; ++k) dst[f * 8 + k] = o[k];
    }
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------

// Window width by log2 size as measured on MI355X by the autotuning sweep
// (batches of 2^(18 - L) MSMs of 2^L points); outside the swept range
// c = L - 3 within [4, 16].
inline int msm_default_window_impl(size_t n) {
    static const int measured[21] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 8, 8, 8, 9, 10, 12, 13, 13, 15, 15, 15};
    int lg = 0;
    while ((size_t(1) << (lg + 1)) <= n) ++lg;
    if (lg >= 10 && lg <= 20) return measured[lg];
    return std::max(4, std::min(16, lg - 3));
}

inline int msm_window_for_impl(size_t n, const WindowTable* table) {
    int lg = 0;
    while ((size_t(1) << (lg + 1)) <= n) ++lg;
    if (table && lg < 25 && table->c[lg] >= 4 && table->c[lg] <= 20) return table->c[lg];
    return msm_default_window_impl(n);
}

inline MsmBatchPlan msm_batch_plan_impl(const std::vector<MsmItem>& items, const MsmConfig& cfg,
                                        const WindowTable* table) {
    require(cfg.chunk >= 1 && cfg.chunk <= 8 && (cfg.chunk & (cfg.chunk - 1)) == 0 && cfg.task_points >= 1 &&
                cfg.task_points <= int(TASK_CLASS_MAX),
            "invalid MSM config");
    MsmBatchPlan p;
    p.task_points = uint32_t(cfg.task_points);
    uint64_t threads = 0, entries = 0, buckets = 0, chunks = 0, windows = 0, hist_blocks = 0, hist_words = 0;
    uint64_t digit_threads = 0, max_window_chunks = 0;
    bool lds = true, group = cfg.task_points <= 32;
    for (const MsmItem& it : items) {
        require(it.n > 0 && it.n <= (size_t(1) << 24), "MSM length must be within 1..2^24");
        const int c = it.c ? it.c : (cfg.window_bits ? cfg.window_bits : msm_window_for_impl(it.n, table));
        require(c >= 4 && c <= 20, "MSM window bits must be within 4..20");
        MsmDesc d{};
        d.n = uint32_t(it.n);
        d.c = uint32_t(c);
        d.windows = uint32_t((255 + c - 1) / c);
        d.digit_windows = d.windows;
        d.wide = wide_windows(d.c, d.windows);
        if (it.table_windows) {
            require(it.table_windows == d.windows, "window-table rows must hold ceil(255 / c) points");
            d.table = 1;
            d.windows = 1;
            d.wide = d.digit_windows;  // the tables hold multiples [2^(c w)] B: every window c bits
        }
        d.buckets = 1u << (c - 1);
        {
            int lg = 0;
            while ((size_t(1) << (lg + 1)) <= it.n) ++lg;
            const int ch = table && lg < 25 && table->chunk[lg] ? table->chunk[lg] : cfg.chunk;
            require(ch >= 1 && ch <= 8 && (ch & (ch - 1)) == 0, "invalid reduction chunk");
            d.chunk = uint32_t(std::min<int>(ch, int(d.buckets)));
        }
        d.point_base = it.point_base;
        d.scalar_base = it.scalar_base;
        d.thread_base = uint32_t(threads);
        d.entry_base = uint32_t(entries);
        d.bucket_base = uint32_t(buckets);
        d.chunk_base = uint32_t(chunks);
        d.window_base = uint32_t(windows);
        d.out_index = uint32_t(p.descs.size());
        d.result_slot = MSM_NO_SLOT;
        d.chunk_shift = 12;
        while (d.chunk_shift < 15 && (1u << d.chunk_shift) < 2 * d.buckets) ++d.chunk_shift;
        d.chunks = uint32_t((it.n + (size_t(1) << d.chunk_shift) - 1) >> d.chunk_shift);
        d.hist_block_base = uint32_t(hist_blocks);
        d.hist_base = uint32_t(hist_words);
        const uint64_t blocks = uint64_t(d.chunks) * (d.table ? 1u : d.digit_windows);
        hist_blocks += blocks;
        hist_words += blocks * d.buckets;
        digit_threads = std::max<uint64_t>(digit_threads, std::min<uint64_t>(it.n, uint64_t(1) << d.chunk_shift));
        lds &= d.buckets <= MSM_LDS_BUCKETS;
        group &= !d.table && d.buckets >= GROUP_BUCKETS;
        threads += it.n;
        entries += uint64_t(d.digit_windows) * it.n;
        buckets += uint64_t(d.windows) * d.buckets;
        chunks += uint64_t(d.windows) * (d.buckets / d.chunk);
        max_window_chunks = std::max<uint64_t>(max_window_chunks, d.buckets / d.chunk);
        windows += d.windows;
        p.descs.push_back(d);
    }
    require(entries < (uint64_t(1) << 31) && buckets < (uint64_t(1) << 31) && hist_words < (uint64_t(1) << 31),
            "MSM batch too large");
    p.total_hist_blocks = uint32_t(hist_blocks);
    p.max_window_chunks = uint32_t(max_window_chunks);
    // Digit blocks need at most one thread per point of their chunk.
    p.digit_threads = uint32_t(std::min<uint64_t>(MSM_DIGIT_THREADS, std::max<uint64_t>(64, (digit_threads + 63) / 64 * 64)));
    p.total_hist = uint32_t(hist_words);
    p.lds_digits = lds && hist_words <= uint64_t(entries) * 2 + (uint64_t(1) << 20);
    // The staging array lives in the partial-sum buffer (at least
    // entries / task_points + 1 XYZZ values of 32 or more words each).
    p.group_place = p.lds_digits && group;
    p.total_threads = uint32_t(threads);
    p.total_entries = uint32_t(entries);
    p.total_buckets = uint32_t(buckets);
    p.total_chunks = uint32_t(chunks);
    p.total_windows = uint32_t(windows);
    const size_t nonempty = std::min<size_t>(entries, buckets);
    p.max_tasks = entries / size_t(cfg.task_points) + nonempty + 1;
    p.max_heavy = std::min<size_t>(buckets, entries / (size_t(cfg.task_points) * HEAVY_TASKS) + 1);
    return p;
}

// Bytes of every device buffer a plan's kernels use (descriptors and the
// window values included), for the growing workspace and for a fixed
// scratch area alike.
struct MsmBufferBytes {
    size_t counts, offsets, digits, ranks, sorted, block_sums, total, tasks, partials, heavy, chunks, hist, descs,
        windows, order;
};

inline MsmBufferBytes msm_buffer_bytes(const MsmBatchPlan& plan, int xw) {
    const size_t tiles = (plan.total_buckets + SCAN_TILE - 1) / SCAN_TILE;
    MsmBufferBytes b;
    b.counts = size_t(plan.total_buckets) * 4;
    b.offsets = size_t(plan.total_buckets) * sizeof(uint2);
    // LDS digits keep no digit array, except the 16-bit planes of group mode
    b.digits = plan.lds_digits ? (plan.group_place ? size_t(plan.total_entries) * 2 : 0)
                               : size_t(plan.total_entries) * 4;
    b.ranks = plan.lds_digits ? 0 : b.digits;
    b.sorted = size_t(plan.total_entries) * 4;
    b.block_sums = tiles * sizeof(uint2);
    b.total = TOTAL_WORDS * 4;
    b.tasks = plan.max_tasks * sizeof(uint4);
    b.partials = plan.max_tasks * size_t(xw) * 4;
    b.heavy = plan.max_heavy * 4;
    b.chunks = std::max<size_t>(size_t(plan.total_chunks) * (plan.chunk_planes ? 2 : 1),
                                size_t(plan.total_planes) * plan.plane_slices) * size_t(xw) * 4;
    b.hist = plan.lds_digits ? size_t(plan.total_hist) * 4 : 0;
    b.descs = plan.descs.size() * sizeof(MsmDesc);
    b.windows = size_t(plan.total_windows) * size_t(xw) * 4;
    b.order = size_t(plan.total_buckets) * 4;
    return b;
}

// The plan's buffers as slices of `capacity` bytes at `base` (256-byte
// aligned); returns the bytes used, or throws when they do not fit.
inline size_t msm_workspace_borrow_impl(const MsmBatchPlan& plan, int xw, uint8_t* base, size_t capacity,
                                        MsmWorkspace& ws) {
    const MsmBufferBytes b = msm_buffer_bytes(plan, xw);
    const std::pair<DeviceBuffer*, size_t> parts[] = {
        {&ws.counts, b.counts},   {&ws.offsets, b.offsets}, {&ws.digits, b.digits},   {&ws.ranks, b.ranks},
        {&ws.sorted, b.sorted},   {&ws.block_sums, b.block_sums}, {&ws.total, b.total}, {&ws.tasks, b.tasks},
        {&ws.partials, b.partials}, {&ws.heavy, b.heavy},   {&ws.chunks, b.chunks},   {&ws.hist, b.hist},
        {&ws.descs, b.descs},     {&ws.windows, b.windows}, {&ws.order, b.order}};
    size_t used = 0;
    for (const auto& p : parts) used += (std::max<size_t>(p.second, 1) + 255) & ~size_t(255);
    require(used <= capacity, "MSM batch needs more scratch than the allocation-free entry point holds");
    size_t at = 0;
    for (const auto& p : parts) {
        const size_t n = (std::max<size_t>(p.second, 1) + 255) & ~size_t(255);
        p.first->borrow(base + at, n);
        at += n;
    }
    return used;
}

inline void msm_batch_prepare_impl(const MsmBatchPlan& plan, int xw, MsmWorkspace& ws, hipStream_t stream) {
    const size_t m = plan.descs.size();
    require(m > 0, "empty MSM batch");
    const MsmBufferBytes b = msm_buffer_bytes(plan, xw);
    ws.counts.ensure(b.counts);
    ws.offsets.ensure(b.offsets);
    if (!plan.lds_digits) {
        ws.digits.ensure(b.digits);
        ws.ranks.ensure(b.ranks);
    } else if (plan.group_place) {
        ws.digits.ensure(b.digits);  // the digit planes
    }
    ws.sorted.ensure(b.sorted);
    ws.block_sums.ensure(b.block_sums);
    ws.total.ensure(b.total);
    ws.tasks.ensure(b.tasks);
    ws.partials.ensure(b.partials);
    ws.heavy.ensure(b.heavy);
    ws.order.ensure(b.order);
    ws.chunks.ensure(b.chunks);
    if (plan.lds_digits) ws.hist.ensure(b.hist);
    MsmDesc* dev = ws.descs.as<MsmDesc>(m);
    // Repeated plans (the same MSM shapes call after call) find their
    // descriptors in place: only kernels read `descs`, and only this
    // function writes it.
    if (ws.uploaded_generation == ws.descs.generation && ws.uploaded.size() == m &&
        std::memcmp(ws.uploaded.data(), plan.descs.data(), m * sizeof(MsmDesc)) == 0)
        return;
    // Earlier work on this stream may still read the pinned staging area.
    HIP_CHECK(hipStreamSynchronize(stream));
    MsmDesc* host = ws.host_descs.as<MsmDesc>(m);
    std::memcpy(host, plan.descs.data(), m * sizeof(MsmDesc));
    ws.uploaded.clear();
    ws.uploaded_generation = ~uint64_t(0);
    HIP_CHECK(hipMemcpyAsync(dev, host, m * sizeof(MsmDesc), hipMemcpyHostToDevice, stream));
    ws.uploaded = plan.descs;
    ws.uploaded_generation = ws.descs.generation;
}

// Buffers of a prepared plan, checked once.
struct MsmBuffers {
    const MsmDesc* descs;
    uint32_t *counts, *digits, *ranks, *sorted, *total, *partials, *heavy, *chunk_vals, *hist, *order;
    uint2 *offsets, *tile_sums;
    uint4* tasks;
    uint32_t m, tiles;
};

inline MsmBuffers msm_buffers(const MsmBatchPlan& plan, MsmWorkspace& ws, int xw) {
    MsmBuffers b;
    b.m = uint32_t(plan.descs.size());
    b.tiles = uint32_t((plan.total_buckets + SCAN_TILE - 1) / SCAN_TILE);
    // The buffers were sized by msm_batch_prepare; check instead of growing.
    require(ws.counts.bytes >= size_t(plan.total_buckets) * 4 && ws.sorted.bytes >= size_t(plan.total_entries) * 4 &&
                ws.total.bytes >= size_t(TOTAL_WORDS) * 4 && ws.tasks.bytes >= plan.max_tasks * sizeof(uint4) &&
                ws.order.bytes >= size_t(plan.total_buckets) * 4 &&
                ws.partials.bytes >= plan.max_tasks * xw * 4 && ws.chunks.bytes >= size_t(plan.total_chunks) * xw * 4 &&
                ws.chunks.bytes >= size_t(plan.total_planes) * plan.plane_slices * xw * 4 &&
                ws.chunks.bytes >= size_t(plan.total_chunks) * (plan.chunk_planes ? 2 : 1) * xw * 4 &&
                ws.descs.bytes >= b.m * sizeof(MsmDesc) && ws.heavy.bytes >= plan.max_heavy * 4 &&
                ws.block_sums.bytes >= b.tiles * sizeof(uint2) &&
                (!plan.lds_digits || ws.hist.bytes >= size_t(plan.total_hist) * 4) &&
                (!plan.group_place || (ws.partials.bytes >= size_t(plan.total_entries) * 4 &&
                                       ws.digits.bytes >= size_t(plan.total_entries) * 2)),
            "MSM workspace not prepared for this batch");
    b.descs = static_cast<const MsmDesc*>(ws.descs.ptr);
    b.counts = static_cast<uint32_t*>(ws.counts.ptr);
    b.offsets = static_cast<uint2*>(ws.offsets.ptr);
    b.digits = static_cast<uint32_t*>(ws.digits.ptr);
    b.ranks = static_cast<uint32_t*>(ws.ranks.ptr);
    b.sorted = static_cast<uint32_t*>(ws.sorted.ptr);
    b.tile_sums = static_cast<uint2*>(ws.block_sums.ptr);
    b.total = static_cast<uint32_t*>(ws.total.ptr);  // TOTAL_WORDS counters (k_tasks)
    b.tasks = static_cast<uint4*>(ws.tasks.ptr);
    b.partials = static_cast<uint32_t*>(ws.partials.ptr);
    b.heavy = static_cast<uint32_t*>(ws.heavy.ptr);
    b.order = static_cast<uint32_t*>(ws.order.ptr);
    b.chunk_vals = static_cast<uint32_t*>(ws.chunks.ptr);
    b.hist = plan.lds_digits ? static_cast<uint32_t*>(ws.hist.ptr) : nullptr;
    return b;
}

inline void msm_mark(hipEvent_t* ev, int k, hipStream_t stream) {
    if (ev) HIP_CHECK(hipEventRecord(ev[k], stream));
}

// Digits, bucket counts, the (count, task) scan, placement and the task list.
// Marks ev[k0 .. k0 + 4]: start, counted, scanned, placed, tasks.
inline void msm_sort_stage(const uint32_t* point_index, const uint32_t* scalars, const MsmBatchPlan& plan,
                           const MsmBuffers& b, hipStream_t stream, hipEvent_t* ev, int k0) {
    const uint32_t m = b.m, tp = plan.task_points;
    const uint32_t per = task_buckets_per_thread(plan.total_buckets);
    const unsigned task_blocks = grid_for(plan.total_buckets, TASK_THREADS * per);
    msm_mark(ev, k0, stream);
    k_fill<<<1, 256, 0, stream>>>(b.total, TOTAL_CLASS_COUNTS + tp, 0u);
    uint16_t* planes = reinterpret_cast<uint16_t*>(b.digits);
    if (plan.group_place) {
        k_digit_planes<<<grid_for(plan.total_threads, 256), 256, 0, stream>>>(b.descs, m, plan.total_threads, scalars,
                                                                             planes);
        k_digits_lds<false, true><<<plan.total_hist_blocks, plan.digit_threads, 0, stream>>>(
            b.descs, m, plan.total_hist_blocks, scalars, b.hist, b.offsets, point_index, plan.point_limit, b.total + 3,
            b.sorted, planes);
        k_hist_prefix<<<grid_for(plan.total_buckets, 256), 256, 0, stream>>>(b.descs, m, plan.total_buckets, b.hist,
                                                                           b.counts, tp, b.total);
    } else if (plan.lds_digits) {
        k_digits_lds<false><<<plan.total_hist_blocks, plan.digit_threads, 0, stream>>>(
            b.descs, m, plan.total_hist_blocks, scalars, b.hist, b.offsets, point_index, plan.point_limit, b.total + 3,
            b.sorted);
        k_hist_prefix<<<grid_for(plan.total_buckets, 256), 256, 0, stream>>>(b.descs, m, plan.total_buckets, b.hist,
                                                                           b.counts, tp, b.total);
    } else {
        k_fill<<<unsigned(std::min<size_t>(grid_for(plan.total_buckets, 256), 4096)), 256, 0, stream>>>(
            b.counts, plan.total_buckets, 0u);
        k_digits<<<grid_for(plan.total_threads, 256), 256, 0, stream>>>(b.descs, m, plan.total_threads, scalars,
                                                                       b.counts, b.digits, b.ranks);
        k_task_classes<<<task_blocks, TASK_THREADS, 0, stream>>>(b.counts, plan.total_buckets, tp, per, b.total);
    }
    HIP_CHECK_LAUNCH();
    msm_mark(ev, k0 + 1, stream);
    k_scan_tiles<<<b.tiles, SCAN_THREADS, 0, stream>>>(b.counts, plan.total_buckets, tp, b.tile_sums);
    k_scan_sums<<<1, SCAN_THREADS, 0, stream>>>(b.tile_sums, b.tiles, tp, b.total);
    k_scan_apply<<<b.tiles, SCAN_THREADS, 0, stream>>>(b.counts, plan.total_buckets, tp, b.tile_sums, b.offsets);
    HIP_CHECK_LAUNCH();
    msm_mark(ev, k0 + 2, stream);
    if (plan.group_place) {
        k_digits_group<<<plan.total_hist_blocks, plan.digit_threads, 0, stream>>>(
            b.descs, m, plan.total_hist_blocks, planes, b.hist, b.offsets, b.partials);
        k_group_place<<<plan.total_buckets / GROUP_BUCKETS, 256, 0, stream>>>(
            b.descs, m, b.counts, b.offsets, b.partials, point_index, plan.point_limit, b.total + 3, b.sorted);
    } else if (plan.lds_digits)
        k_digits_lds<true><<<plan.total_hist_blocks, plan.digit_threads, 0, stream>>>(
            b.descs, m, plan.total_hist_blocks, scalars, b.hist, b.offsets, point_index, plan.point_limit, b.total + 3,
            b.sorted);
    else
        k_scatter<<<grid_for(plan.total_entries, 256), 256, 0, stream>>>(b.descs, m, plan.total_entries, b.digits,
                                                                        b.ranks, b.offsets, point_index,
                                                                        plan.point_limit, b.total + 3, b.sorted);
    HIP_CHECK_LAUNCH();
    msm_mark(ev, k0 + 3, stream);
    k_tasks<<<task_blocks, TASK_THREADS, 0, stream>>>(b.counts, b.offsets, plan.total_buckets, tp, per, b.tasks,
                                                      b.heavy, b.order, b.total);
    HIP_CHECK_LAUNCH();
    msm_mark(ev, k0 + 4, stream);
}

template <class F>
void msm_accumulate_stage(const uint32_t* points, const MsmBatchPlan& plan, const MsmBuffers& b,
                          hipStream_t stream) {
    k_accumulate<F><<<grid_for(plan.max_tasks, 256), 256, 0, stream>>>(points, b.sorted, b.tasks, b.total,
                                                                        b.partials);
    HIP_CHECK_LAUNCH();
}

// Heavy buckets, bucket sums and the window reduction.  Marks ev[k0] after
// the bucket sums and ev[k0 + 1] after the window values.
template <class F>
void msm_reduce_stage(const MsmBatchPlan& plan, const MsmBuffers& b, hipStream_t stream, uint32_t* window_out,
                      hipEvent_t* ev, int k0) {
    const uint32_t m = b.m, tp = plan.task_points;
    k_heavy<F><<<unsigned(std::min<size_t>(plan.max_heavy, 1024)), 256, 0, stream>>>(b.heavy, b.total + 2, b.counts,
                                                                                     b.offsets, tp, b.partials);
    k_bucket_sums<F><<<grid_for(plan.total_buckets, 256), 256, 0, stream>>>(b.counts, b.offsets, b.order, b.total,
                                                                           tp, b.partials);
    if (plan.planes) {
        k_plane_sums<F><<<plan.total_planes * plan.plane_slices, PLANE_THREADS, 0, stream>>>(
            b.descs, m, plan.plane_slices, b.counts, b.offsets, b.partials, b.chunk_vals);
        HIP_CHECK_LAUNCH();
        msm_mark(ev, k0, stream);
        k_plane_final<F><<<grid_for(plan.total_planes, 64), 64, 0, stream>>>(b.chunk_vals, plan.total_planes,
                                                                               plan.plane_slices, window_out);
        HIP_CHECK_LAUNCH();
        msm_mark(ev, k0 + 1, stream);
        return;
    }
    if (plan.chunk_planes) {
        k_reduce_chunks<F, true><<<grid_for(plan.total_chunks, 256), 256, 0, stream>>>(
            b.descs, m, plan.total_chunks, b.counts, b.offsets, b.partials, b.chunk_vals);
        HIP_CHECK_LAUNCH();
        msm_mark(ev, k0, stream);
        k_chunk_planes<F><<<plan.total_planes, PLANE_THREADS, 0, stream>>>(b.descs, m, b.chunk_vals, window_out);
        HIP_CHECK_LAUNCH();
        msm_mark(ev, k0 + 1, stream);
        return;
    }
    k_reduce_chunks<F><<<grid_for(plan.total_chunks, 256), 256, 0, stream>>>(
        b.descs, m, plan.total_chunks, b.counts, b.offsets, b.partials, b.chunk_vals);
    HIP_CHECK_LAUNCH();
    msm_mark(ev, k0, stream);
    if (plan.max_window_chunks <= 32)  // small bucket sets: one value per thread, fewest levels
        k_window_tree<F, 32><<<plan.total_windows, 32, 0, stream>>>(b.descs, m, b.chunk_vals, window_out);
    else if (plan.max_window_chunks <= 128)
        k_window_tree<F, 128><<<plan.total_windows, 128, 0, stream>>>(b.descs, m, b.chunk_vals, window_out);
    else
        k_window_tree<F, TreeThreads<F>::value><<<plan.total_windows, TreeThreads<F>::value, 0, stream>>>(
            b.descs, m, b.chunk_vals, window_out);
    HIP_CHECK_LAUNCH();
    msm_mark(ev, k0 + 1, stream);
}

template <class F>
void msm_batch_enqueue(const uint32_t* points, const uint32_t* point_index, const uint32_t* scalars,
                       const MsmBatchPlan& plan, MsmWorkspace& ws, hipStream_t stream, uint32_t* window_out) {
    const MsmBuffers b = msm_buffers(plan, ws, PointLayout<F>::XYZZ_WORDS);
    hipEvent_t* ev = ws.phase_events;  // start, digits, scan, scatter, tasks, accumulate, reduce, tree
    msm_sort_stage(point_index, scalars, plan, b, stream, ev, 0);
    msm_accumulate_stage<F>(points, plan, b, stream);
    msm_mark(ev, 5, stream);
    msm_reduce_stage<F>(plan, b, stream, window_out, ev, 6);
}

template <class F>
void msm_batch_finalize_enqueue(const uint32_t* window_words, const MsmBatchPlan& plan, const MsmWorkspace& ws,
                                uint8_t* out, uint32_t* slot_points, hipStream_t stream) {
    const uint32_t m = uint32_t(plan.descs.size());
    require(ws.descs.bytes >= m * sizeof(MsmDesc), "MSM workspace not prepared for this batch");
    k_finalize<F><<<grid_for(m, 64), 64, 0, stream>>>(static_cast<const MsmDesc*>(ws.descs.ptr), m, window_words,
                                                      out, slot_points, static_cast<const uint32_t*>(ws.total.ptr) + 3);
    HIP_CHECK_LAUNCH();
}

// Point arithmetic of the host epilogues, in the host's own 64-bit-limb
// field: G1 (host_g1.hpp) and G2 (host_g2.hpp).
template <class F>
struct HostArith;
template <>
struct HostArith<FqOps> {
    using P = hostg1::Xyzz;
    static P load(const uint32_t* w) { return hostg1::from_device_xyzz(w); }
    static P inf() { return hostg1::infinity(); }
    static P add(const P& a, const P& b) { return hostg1::add(a, b); }
    static P dbl(const P& a) { return hostg1::dbl(a); }
    static P dbl_n(const P& a, int n) { return hostg1::dbl_n(a, n); }
    static Affine<FqOps> affine(const P& p) {
        if (hostg1::is_inf(p)) return {fp_zero<FqParams>(), fp_zero<FqParams>(), true};
        uint32_t x[8], y[8];
        hostg1::to_affine_words(p, x, y);
        return {to_mont<FqParams>(x), to_mont<FqParams>(y), false};
    }
};
template <>
struct HostArith<Fq2Ops> {
    using P = hostg2::Xyzz;
    static P load(const uint32_t* w) { return hostg2::from_device_xyzz(w); }
    static P inf() { return hostg2::infinity(); }
    static P add(const P& a, const P& b) { return hostg2::add(a, b); }
    static P dbl(const P& a) { return hostg2::dbl(a); }
    static P dbl_n(const P& a, int n) { return hostg2::dbl_n(a, n); }
    static Affine<Fq2Ops> affine(const P& p) {
        if (hostg2::is_inf(p)) return {f2_zero(), f2_zero(), true};
        uint32_t w[4][8];
        hostg2::to_affine_words(p, w);
        return {{to_mont<FqParams>(w[0]), to_mont<FqParams>(w[1])}, {to_mont<FqParams>(w[2]), to_mont<FqParams>(w[3])},
                false};
    }
};

template <class F>
Affine<F> combine_one(const uint32_t* words, const MsmDesc& d) {
    using H = HostArith<F>;
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    const uint32_t* base = words + size_t(d.window_base) * XW;
    typename H::P acc = H::load(base + size_t(d.windows - 1) * XW);
    for (int w = int(d.windows) - 2; w >= 0; --w) {
        acc = H::add(H::dbl_n(acc, int(window_width(d, uint32_t(w)))), H::load(base + size_t(w) * XW));
    }
    return H::affine(acc);
}

// Window combination on the host, one MSM per pool task (a CPU core runs
// the ~250 serial doublings of one MSM about ten times faster than a GPU lane).
template <class F>
void msm_batch_combine(const uint32_t* words, const MsmBatchPlan& plan, Affine<F>* out) {
    HostPool::get().parallel_for(plan.descs.size(),
                                 [&](size_t i) { out[i] = combine_one<F>(words, plan.descs[i]); });
}

// ---- single-MSM conveniences -------------------------------------------------

inline MsmPlan msm_plan_impl(size_t n, const MsmConfig& cfg) {
    MsmPlan p;
    p.c = cfg.window_bits ? cfg.window_bits : msm_default_window_impl(n);
    p.windows = (255 + p.c - 1) / p.c;
    p.wide = int(wide_windows(uint32_t(p.c), uint32_t(p.windows)));
    p.buckets = 1u << (p.c - 1);
    p.n = n;
    p.planes = int(plane_bits(p.buckets));
    // room for either reduction's output (chunk planes need the most with
    // one bucket per chunk)
    p.outputs = size_t(p.windows) * std::max<size_t>(size_t(p.planes), chunk_plane_shape(p.buckets, 1).per_window);
    return p;
}

// The plan of one MSM (or one window range of it) whose windows are
// combined on the host: bit planes of the bucket weights for small bucket
// sets, of the chunk index otherwise.
inline MsmBatchPlan msm_single_plan_impl(const MsmItem& item, MsmConfig c, bool g2) {
    if (g2) c.chunk = std::min(c.chunk, 2);  // G2: short reduction chains
    MsmBatchPlan plan = msm_batch_plan_impl({item}, c, nullptr);
    if (!msm_plan_planes(plan)) msm_plan_chunk_planes(plan);
    return plan;
}

// What a single-MSM plan leaves per window, for the host combination.
inline MsmPlan msm_outputs_of(const MsmBatchPlan& plan) {
    MsmPlan p;
    p.c = int(plan.descs[0].c);
    p.windows = int(plan.descs[0].windows);
    p.wide = int(plan.descs[0].wide);
    p.buckets = plan.descs[0].buckets;
    p.n = plan.descs[0].n;
    p.planes = plan.planes ? int(plane_bits(p.buckets))
                           : int(chunk_plane_shape(p.buckets, plan.descs[0].chunk).per_window);
    p.chunk = plan.chunk_planes ? int(plan.descs[0].chunk) : 0;
    p.outputs = plan.total_planes;
    return p;
}

template <class F>
MsmPlan msm_enqueue(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                    hipStream_t stream, const MsmConfig& cfg, uint32_t* window_out, size_t out_capacity) {
    MsmItem item{n, 0, 0, cfg.window_bits};
    const MsmBatchPlan plan = msm_single_plan_impl(item, cfg, std::is_same<F, Fq2Ops>::value);
    require(plan.total_planes <= out_capacity, "MSM window buffer too small");
    msm_batch_prepare_impl(plan, PointLayout<F>::XYZZ_WORDS, ws, stream);
    msm_batch_enqueue<F>(points, nullptr, scalars, plan, ws, stream, window_out);
    return msm_outputs_of(plan);
}

// words: plan.outputs values from msm_enqueue — one per window, or in plane
// mode (plan.planes > 0) per window the bit-plane sums (Horner over the
// planes) or, in chunk-plane mode, U + L * sum_k 2^k Q_k from
// [U slices][Q_0 slices][Q_1 slices]...
template <class F>
typename HostArith<F>::P msm_window_value(const uint32_t* words, const MsmPlan& plan, size_t w) {
    using H = HostArith<F>;
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    const uint32_t* base = words + w * size_t(plan.planes) * XW;
    if (plan.chunk) {
        const ChunkPlaneShape sh = chunk_plane_shape(plan.buckets, uint32_t(plan.chunk));
        auto slices = [&](size_t first, uint32_t count) {
            typename H::P v = H::load(base + first * XW);
            for (uint32_t q = 1; q < count; ++q) v = H::add(v, H::load(base + (first + q) * XW));
            return v;
        };
        typename H::P acc = H::inf();
        for (int k = int(sh.qb) - 1; k >= 0; --k) acc = H::add(H::dbl(acc), slices(sh.su + size_t(k) * sh.sq, sh.sq));
        int shift = 0;
        for (int l = plan.chunk; l > 1; l >>= 1) ++shift;
        if (shift) acc = H::dbl_n(acc, shift);
        return H::add(acc, slices(0, sh.su));
    }
    typename H::P acc = H::load(base + size_t(plan.planes - 1) * XW);
    for (int k = plan.planes - 2; k >= 0; --k) acc = H::add(H::dbl(acc), H::load(base + size_t(k) * XW));
    return acc;
}

// Horner over the windows, sum_w 2^(w c) V_w.  The window values are
// computed on the host pool, top window first, while the calling thread
// runs the chain of doublings and takes each value as it needs it.
template <class F>
Affine<F> msm_combine_windows(const uint32_t* words, const MsmPlan& plan) {
    if (!plan.planes) {
        MsmDesc d{};
        d.c = uint32_t(plan.c);
        d.windows = uint32_t(plan.windows);
        d.wide = uint32_t(plan.wide);
        d.window_base = 0;
        return combine_one<F>(words, d);
    }
    using H = HostArith<F>;
    require(plan.outputs == size_t(plan.windows) * size_t(plan.planes), "MSM plane count mismatch");
    const size_t W = size_t(plan.windows);
    std::vector<typename H::P> win(W);
    std::unique_ptr<std::atomic<int>[]> ready(new std::atomic<int>[W]);
    for (size_t w = 0; w < W; ++w) ready[w].store(0, std::memory_order_relaxed);
    auto one = [&](size_t k) {  // job index k is window W - 1 - k
        const size_t w = W - 1 - k;
        win[w] = msm_window_value<F>(words, plan, w);
        ready[w].store(1, std::memory_order_release);
    };
    const auto job = HostPool::get().start(W, one);
    auto wait_for = [&](size_t w) {
        while (!ready[w].load(std::memory_order_acquire))
            if (!HostPool::help(*job)) std::this_thread::yield();
    };
    wait_for(W - 1);
    typename H::P acc = win[W - 1];
    for (size_t w = W - 1; w-- > 0;) {
        acc = H::dbl_n(acc, int(w) < plan.wide ? plan.c : plan.c - 1);
        wait_for(w);
        acc = H::add(acc, win[w]);
    }
    HostPool::finish(*job);
    return H::affine(acc);
}

template <class F>
Affine<F> msm_run(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                  hipStream_t stream, const MsmConfig& cfg) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    const MsmPlan plan0 = msm_plan_impl(n, cfg);
    uint32_t* window_words = ws.windows.as<uint32_t>(plan0.outputs * XW);
    const MsmPlan plan = msm_enqueue<F>(points, scalars, n, ws, stream, cfg, window_words, plan0.outputs);
    uint32_t* host = ws.host_windows.as<uint32_t>(plan.outputs * XW);
    HIP_CHECK(hipMemcpyAsync(host, window_words, plan.outputs * XW * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    return msm_combine_windows<F>(host, plan);
}

template <class F>
void points_to_mont(const uint8_t* canonical, uint32_t* mont, size_t n, hipStream_t stream) {
    if (!n) return;
    k_to_mont<F><<<grid_for(n, 128), 128, 0, stream>>>(canonical, mont, uint32_t(n));
    HIP_CHECK_LAUNCH();
}

template <class F>
void affine_to_bytes(const Affine<F>& a, uint8_t* out) {
    constexpr int B = PointLayout<F>::CANON_BYTES;
    std::memset(out, 0, B);
    if (a.inf) return;
    auto put = [&](const Fq& v, int slot) {
        uint32_t w[8];
        from_mont(v, w);
        for (int k = 0; k < 8; ++k)
            for (int b = 0; b < 4; ++b) out[slot * 32 + 4 * k + b] = uint8_t(w[k] >> (8 * b));
    };
    if constexpr (B == 64) {
        put(a.x, 0);
        put(a.y, 1);
    } else {
        put(a.x.c0, 0);
        put(a.x.c1, 1);
        put(a.y.c0, 2);
        put(a.y.c1, 3);
    }
}

}  // namespace dvgpu
