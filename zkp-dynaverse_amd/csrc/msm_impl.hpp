// Pippenger MSM for BN254 G1/G2 on gfx950.
//
// Phases (one stream, no host synchronisation until the window values):
//   1. digits   : scalar -> ceil(255/c) signed c-bit digits; each nonzero digit
//                 takes a rank in its bucket with one global atomic.
//   2. scan     : exclusive scan of (bucket count, task count) pairs.
//   3. scatter  : point index (plus sign bit) to its bucket-sorted slot.
//   4. tasks    : every bucket is cut into tasks of at most T points, so one
//                 hot bucket (repeated scalars) spreads over many threads.
//   5. accumulate: one thread per task, XYZZ mixed additions -> partial sums.
//   6. reduce   : one thread per chunk of L buckets computes
//                 sum_i (i+1) S_i by running sums and adds (jL) * sum_i S_i.
//   7. tree     : per-window tree sum of the chunk values.
// The W window values go to the host, which forms sum_w 2^(wc) V_w: ~250
// serial doublings that one CPU core finishes far sooner than one GPU lane.
#pragma once
#include "msm.hpp"

#include <algorithm>
#include <cstring>

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

// bits [bit, bit + c) of a 256-bit little-endian scalar (c <= 24)
__device__ __forceinline__ uint32_t scalar_bits(const uint32_t (&s)[8], int bit, int c) {
    const int j = bit >> 5, sh = bit & 31;
    uint64_t v = 0;
    if (j < 8) v = s[j];
    if (j + 1 < 8) v |= uint64_t(s[j + 1]) << 32;
    return uint32_t(v >> sh) & ((1u << c) - 1u);
}

static __global__ void k_digits(const uint32_t* __restrict__ scalars, uint32_t n, int c, int windows,
                         uint32_t buckets, uint32_t* __restrict__ counts,
                         uint32_t* __restrict__ digits, uint32_t* __restrict__ ranks) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4* sp = reinterpret_cast<const uint4*>(scalars + size_t(i) * 8);
    const uint4 a = sp[0], b = sp[1];
    const uint32_t s[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t carry = 0;
    const uint32_t full = 1u << c;
    for (int w = 0; w < windows; ++w) {
        uint32_t d = scalar_bits(s, w * c, c) + carry;
        uint32_t sign = 0;
        if (d > buckets) {
            d = full - d;
            sign = 1;
            carry = 1;
        } else {
            carry = 0;
        }
        const size_t slot = size_t(w) * n + i;
        if (d == 0) {
            digits[slot] = NO_DIGIT;
        } else {
            const uint32_t bucket = d - 1;
            ranks[slot] = atomicAdd(&counts[size_t(w) * buckets + bucket], 1u);
            digits[slot] = (bucket << 1) | sign;
        }
    }
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
// total (points, tasks) to total[0..1].
static __global__ void k_scan_sums(uint2* __restrict__ tile_sums, uint32_t tiles, uint32_t* __restrict__ total) {
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

static __global__ void k_scatter(const uint32_t* __restrict__ digits, const uint32_t* __restrict__ ranks,
                          const uint2* __restrict__ offsets, uint32_t n, int windows, uint32_t buckets,
                          uint32_t* __restrict__ sorted) {
    const size_t slot = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (slot >= size_t(windows) * n) return;
    const uint32_t d = digits[slot];
    if (d == NO_DIGIT) return;
    const uint32_t w = uint32_t(slot / n), i = uint32_t(slot % n);
    const uint32_t pos = offsets[size_t(w) * buckets + (d >> 1)].x + ranks[slot];
    sorted[pos] = i | ((d & 1u) << 31);
}

// One thread per bucket: emit (start, length) of each of its tasks.
// Buckets cut into more than HEAVY_TASKS tasks have their partial sums
// combined by a whole block (k_heavy) instead of serially in k_reduce_chunks.
constexpr uint32_t HEAVY_TASKS = 8;

static __global__ void k_tasks(const uint32_t* __restrict__ counts, const uint2* __restrict__ offsets,
                               uint32_t m, uint32_t task_points, uint2* __restrict__ tasks,
                               uint32_t* __restrict__ heavy, uint32_t* __restrict__ heavy_count) {
    const uint32_t wb = blockIdx.x * blockDim.x + threadIdx.x;
    if (wb >= m) return;
    const uint32_t cnt = counts[wb];
    const uint2 off = offsets[wb];
    uint32_t k = 0;
    for (uint32_t s = 0; s < cnt; ++k, s += task_points)
        tasks[off.y + k] = make_uint2(off.x + s, min(task_points, cnt - s));
    if (k > HEAVY_TASKS) heavy[atomicAdd(heavy_count, 1u)] = wb;
}

template <class F>
__global__ void __launch_bounds__(256) k_accumulate(const uint32_t* __restrict__ points,
                                                    const uint32_t* __restrict__ sorted,
                                                    const uint2* __restrict__ tasks,
                                                    const uint32_t* __restrict__ total,
                                                    uint32_t* __restrict__ partials) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total[1]) return;
    const uint2 task = tasks[t];
    Xyzz<F> acc = xyzz_inf<F>();
    for (uint32_t e = 0; e < task.y; ++e) {
        const uint32_t entry = sorted[task.x + e];
        const uint32_t idx = entry & 0x7fffffffu;
        Affine<F> p = Mem<F>::load_affine(points + size_t(idx) * PointLayout<F>::AFFINE_WORDS);
        if (p.inf) continue;
        if (entry >> 31) p.y = F::neg(p.y);
        acc = xyzz_madd<F>(acc, p.x, p.y);
    }
    Mem<F>::store_xyzz(partials + size_t(t) * PointLayout<F>::XYZZ_WORDS, acc);
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

// One thread per chunk of L buckets of one window:
//   out = sum_i (i + 1) S_{jL+i} + (jL) * sum_i S_{jL+i}
// where S_b is the sum of bucket b's task partials.
template <class F>
__global__ void __launch_bounds__(256) k_reduce_chunks(const uint32_t* __restrict__ counts,
                                                       const uint2* __restrict__ offsets,
                                                       const uint32_t* __restrict__ partials,
                                                       uint32_t buckets, int windows, int chunk,
                                                       uint32_t task_points,
                                                       uint32_t* __restrict__ out) {
    const uint32_t per_window = buckets / chunk;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= per_window * uint32_t(windows)) return;
    const uint32_t w = g / per_window, j = g % per_window;
    Xyzz<F> run = xyzz_inf<F>(), acc = xyzz_inf<F>();
    for (int i = chunk - 1; i >= 0; --i) {
        const size_t wb = size_t(w) * buckets + size_t(j) * chunk + i;
        const uint32_t cnt = counts[wb];
        if (cnt) {
            const uint32_t t0 = offsets[wb].y;
            uint32_t nt = (cnt + task_points - 1) / task_points;
            if (nt > HEAVY_TASKS) nt = 1;  // combined by k_heavy
            Xyzz<F> s = Mem<F>::load_xyzz(partials + size_t(t0) * PointLayout<F>::XYZZ_WORDS);
            for (uint32_t k = 1; k < nt; ++k)
                s = xyzz_add<F>(s, Mem<F>::load_xyzz(partials + size_t(t0 + k) * PointLayout<F>::XYZZ_WORDS));
            run = xyzz_add<F>(run, s);
        }
        acc = xyzz_add<F>(acc, run);
    }
    const uint32_t offset = j * uint32_t(chunk);
    if (offset) {
        const uint32_t k[1] = {offset};
        acc = xyzz_add<F>(acc, xyzz_mul_words<F, 1>(run, k, 32 - __clz(offset)));
    }
    Mem<F>::store_xyzz(out + size_t(g) * PointLayout<F>::XYZZ_WORDS, acc);
}

// Per-window tree level: out[w][k] = sum of in[w][k*fan .. k*fan + fan).
template <class F>
__global__ void __launch_bounds__(256) k_tree(const uint32_t* __restrict__ in, uint32_t in_per_window,
                                              int windows, int fan, uint32_t* __restrict__ out) {
    const uint32_t out_per_window = (in_per_window + fan - 1) / fan;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= out_per_window * uint32_t(windows)) return;
    const uint32_t w = g / out_per_window, k = g % out_per_window;
    const uint32_t first = k * fan, last = min(first + fan, in_per_window);
    const size_t base = size_t(w) * in_per_window;
    Xyzz<F> acc = Mem<F>::load_xyzz(in + (base + first) * PointLayout<F>::XYZZ_WORDS);
    for (uint32_t e = first + 1; e < last; ++e)
        acc = xyzz_add<F>(acc, Mem<F>::load_xyzz(in + (base + e) * PointLayout<F>::XYZZ_WORDS));
    Mem<F>::store_xyzz(out + size_t(g) * PointLayout<F>::XYZZ_WORDS, acc);
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

inline int msm_default_window_impl(size_t n) {
    int lg = 0;
    while ((size_t(1) << (lg + 1)) <= n) ++lg;
    int c = lg - 3;
    return std::max(4, std::min(16, c));
}

inline MsmPlan msm_plan_impl(size_t n, const MsmConfig& cfg) {
    MsmPlan p;
    p.c = cfg.window_bits ? cfg.window_bits : msm_default_window_impl(n);
    require(p.c >= 4 && p.c <= 20, "MSM window bits must be within 4..20");
    p.windows = (255 + p.c - 1) / p.c;
    p.buckets = 1u << (p.c - 1);
    p.n = n;
    const size_t entries = size_t(p.windows) * n;
    const size_t nonempty = std::min<size_t>(entries, size_t(p.windows) * p.buckets);
    p.max_tasks = entries / size_t(cfg.task_points) + nonempty + 1;
    return p;
}

template <class F>
MsmPlan msm_enqueue(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                    hipStream_t stream, const MsmConfig& cfg, uint32_t* window_out) {
    require(n > 0 && n <= (size_t(1) << 24), "MSM length must be within 1..2^24");
    require(cfg.chunk >= 1 && cfg.task_points >= 1 && cfg.tree_fanin >= 2, "invalid MSM config");
    const MsmPlan plan = msm_plan_impl(n, cfg);
    require(plan.buckets % uint32_t(cfg.chunk) == 0, "chunk must divide bucket count");
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    const int W = plan.windows;
    const uint32_t B = plan.buckets;
    const size_t m = size_t(W) * B;
    const size_t entries = size_t(W) * n;
    const uint32_t tiles = uint32_t((m + SCAN_TILE - 1) / SCAN_TILE);

    uint32_t* counts = ws.counts.as<uint32_t>(m);
    uint2* offsets = ws.offsets.as<uint2>(m);
    uint32_t* digits = ws.digits.as<uint32_t>(entries);
    uint32_t* ranks = ws.ranks.as<uint32_t>(entries);
    uint32_t* sorted = ws.sorted.as<uint32_t>(entries);
    uint2* tile_sums = ws.block_sums.as<uint2>(tiles);
    uint32_t* total = ws.total.as<uint32_t>(4);  // [0] entries, [1] tasks, [2] heavy buckets
    uint2* tasks = ws.tasks.as<uint2>(plan.max_tasks);
    uint32_t* partials = ws.partials.as<uint32_t>(plan.max_tasks * XW);
    const uint32_t chunks_per_window = B / uint32_t(cfg.chunk);
    uint32_t* chunk_vals = ws.chunks.as<uint32_t>(size_t(chunks_per_window) * W * XW);

    hipEvent_t* ev = ws.phase_events;
    auto mark = [&](int k) {
        if (ev) HIP_CHECK(hipEventRecord(ev[k], stream));
    };
    mark(0);
    const size_t max_heavy = std::min<size_t>(m, entries / (size_t(cfg.task_points) * HEAVY_TASKS) + 1);
    uint32_t* heavy = ws.heavy.as<uint32_t>(max_heavy);
    HIP_CHECK(hipMemsetAsync(counts, 0, m * sizeof(uint32_t), stream));
    HIP_CHECK(hipMemsetAsync(total, 0, 4 * sizeof(uint32_t), stream));
    k_digits<<<grid_for(n, 256), 256, 0, stream>>>(scalars, uint32_t(n), plan.c, W, B, counts, digits, ranks);
    HIP_CHECK_LAUNCH();
    mark(1);
    const uint32_t tp = uint32_t(cfg.task_points);
    k_scan_tiles<<<tiles, SCAN_THREADS, 0, stream>>>(counts, uint32_t(m), tp, tile_sums);
    HIP_CHECK_LAUNCH();
    k_scan_sums<<<1, SCAN_THREADS, 0, stream>>>(tile_sums, tiles, total);
    HIP_CHECK_LAUNCH();
    k_scan_apply<<<tiles, SCAN_THREADS, 0, stream>>>(counts, uint32_t(m), tp, tile_sums, offsets);
    HIP_CHECK_LAUNCH();
    mark(2);
    k_scatter<<<grid_for(entries, 256), 256, 0, stream>>>(digits, ranks, offsets, uint32_t(n), W, B, sorted);
    HIP_CHECK_LAUNCH();
    mark(3);
    k_tasks<<<grid_for(m, 256), 256, 0, stream>>>(counts, offsets, uint32_t(m), tp, tasks, heavy, total + 2);
    HIP_CHECK_LAUNCH();
    mark(4);
    k_accumulate<F><<<grid_for(plan.max_tasks, 256), 256, 0, stream>>>(points, sorted, tasks, total, partials);
    HIP_CHECK_LAUNCH();
    k_heavy<F><<<unsigned(std::min<size_t>(max_heavy, 1024)), 256, 0, stream>>>(heavy, total + 2, counts, offsets,
                                                                               tp, partials);
    HIP_CHECK_LAUNCH();
    mark(5);
    const size_t nchunks = size_t(chunks_per_window) * W;
    k_reduce_chunks<F><<<grid_for(nchunks, 256), 256, 0, stream>>>(counts, offsets, partials, B, W,
                                                                   cfg.chunk, tp, chunk_vals);
    HIP_CHECK_LAUNCH();
    mark(6);
    // Tree levels down to one value per window.
    uint32_t per_window = chunks_per_window;
    const uint32_t* src = chunk_vals;
    int flip = 0;
    while (per_window > 1) {
        const uint32_t next = (per_window + cfg.tree_fanin - 1) / cfg.tree_fanin;
        uint32_t* dst = next == 1 ? window_out
                                  : (flip ? ws.tree_b.as<uint32_t>(size_t(next) * W * XW)
                                          : ws.tree_a.as<uint32_t>(size_t(next) * W * XW));
        k_tree<F><<<grid_for(size_t(next) * W, 256), 256, 0, stream>>>(src, per_window, W, cfg.tree_fanin, dst);
        HIP_CHECK_LAUNCH();
        src = dst;
        per_window = next;
        flip ^= 1;
    }
    if (src != window_out)
        HIP_CHECK(hipMemcpyAsync(window_out, src, size_t(W) * XW * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, stream));
    mark(7);
    return plan;
}

template <class F>
struct HostMem;
template <>
struct HostMem<FqOps> {
    static Fq ld(const uint32_t* p) {
        uint32_t w[8];
        std::memcpy(w, p, 32);
        return unpack<FqParams>(w);
    }
    static Xyzz<FqOps> load_xyzz(const uint32_t* p) { return {ld(p), ld(p + 8), ld(p + 16), ld(p + 24)}; }
};
template <>
struct HostMem<Fq2Ops> {
    static Fq2 ld(const uint32_t* p) { return {HostMem<FqOps>::ld(p), HostMem<FqOps>::ld(p + 8)}; }
    static Xyzz<Fq2Ops> load_xyzz(const uint32_t* p) { return {ld(p), ld(p + 16), ld(p + 32), ld(p + 48)}; }
};

template <class F>
Affine<F> msm_combine_windows(const uint32_t* words, const MsmPlan& plan) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    Xyzz<F> acc = HostMem<F>::load_xyzz(words + size_t(plan.windows - 1) * XW);
    for (int w = plan.windows - 2; w >= 0; --w) {
        for (int k = 0; k < plan.c; ++k) acc = xyzz_dbl<F>(acc);
        acc = xyzz_add<F>(acc, HostMem<F>::load_xyzz(words + size_t(w) * XW));
    }
    return xyzz_to_affine<F>(acc);
}

template <class F>
Affine<F> msm_run(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                  hipStream_t stream, const MsmConfig& cfg) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    const MsmPlan plan0 = msm_plan_impl(n, cfg);
    uint32_t* window_words = ws.windows.as<uint32_t>(size_t(plan0.windows) * XW);
    const MsmPlan plan = msm_enqueue<F>(points, scalars, n, ws, stream, cfg, window_words);
    uint32_t* host = ws.host_windows.as<uint32_t>(size_t(plan.windows) * XW);
    HIP_CHECK(hipMemcpyAsync(host, window_words, size_t(plan.windows) * XW * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, stream));
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
