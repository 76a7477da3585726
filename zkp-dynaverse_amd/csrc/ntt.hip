// Natural-order radix-2 NTT over BN254 Fr (decimation in frequency).
//
// A transform of length 2^L runs as P = ceil(L / 11) LDS passes plus one
// permutation pass.  A pass owns a bit range [a, a + S) of the element index
// and runs those S butterfly stages on tiles of 2^S x C elements held in LDS
// (C consecutive low indices keep the global accesses contiguous).  The
// butterfly is (x, y) -> (x + y, (x - y) * w): x - y is formed limb-wise
// without a carry pass and fed straight into the Montgomery product.  The last
// pass is followed by a bit-reversal gather that also leaves Montgomery form
// (and applies 1/n on the inverse) with one product per element.
#include "ntt.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace dvgpu {

using namespace bn254;

namespace {

constexpr int TILE = 2048;  // elements per LDS tile (72 KiB as nine limb planes)
constexpr int THREADS = 256;        // element-wise kernels
constexpr int PASS_THREADS = 1024;  // DIF passes: one butterfly per thread and stage on a full tile

// Threads for a DIF pass over tiles of `tile` elements (one per butterfly).
inline unsigned pass_threads(uint32_t tile) {
    return unsigned(std::min<uint32_t>(PASS_THREADS, std::max<uint32_t>(64, tile / 2)));
}
constexpr int MAX_STAGES = 11;

__device__ __forceinline__ Fr ld8(const uint32_t* p) {
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return unpack<FrParams>(w);
}

__device__ __forceinline__ void st8(uint32_t* p, const Fr& v) {
    uint32_t w[8];
    pack(v, w);
    reinterpret_cast<uint4*>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// One DIF pass over index bits [a, a + S) of `rows` transforms of length 2^L;
// `g` is the tile of this block.  CANON_IN: the input holds canonical
// integers (converted on load).
// The tile lives in LDS as nine limb planes (limb k of element e at
// lds[k * TILE + e]): butterflies read and write limbs directly, with no
// packing, and consecutive lanes touch consecutive words.
__device__ __forceinline__ Fr lds_ld(const uint32_t* __restrict__ lds, uint32_t e) {
    Fr x;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) x.l[k] = lds[k * TILE + e];
    return x;
}

__device__ __forceinline__ void lds_st(uint32_t* __restrict__ lds, uint32_t e, const Fr& x) {
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) lds[k * TILE + e] = x.l[k];
}

template <bool CANON_IN>
__device__ __forceinline__ void dif_tile(uint32_t* __restrict__ lds, const uint32_t* __restrict__ in,
                                         uint32_t* __restrict__ out, int L, int a, int S, int C,
                                         const uint32_t* __restrict__ tw, size_t g) {
    const size_t n = size_t(1) << L;
    const uint32_t tile_elems = (1u << S) * C;
    const size_t tiles_per_row = n / tile_elems;
    const size_t row = g / tiles_per_row;
    const size_t t = g % tiles_per_row;
    const size_t lo_groups = (size_t(1) << a) / C;
    const size_t hi = t / lo_groups;
    const size_t lo0 = (t % lo_groups) * C;
    const size_t base = row * n + (hi << (a + S)) + lo0;

    for (uint32_t e = threadIdx.x; e < tile_elems; e += blockDim.x) {
        const uint32_t mid = e / C, c = e % C;
        const uint32_t* src = in + (base + (size_t(mid) << a) + c) * 8;
        const uint4 v0 = reinterpret_cast<const uint4*>(src)[0];
        const uint4 v1 = reinterpret_cast<const uint4*>(src)[1];
        const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        lds_st(lds, e, CANON_IN ? to_mont<FrParams>(w) : unpack<FrParams>(w));
    }
    __syncthreads();

    const uint32_t half = tile_elems / 2;
    for (int sl = S - 1; sl >= 0; --sl) {
        const int s = a + sl;
        const uint32_t h = 1u << sl;
        for (uint32_t b = threadIdx.x; b < half; b += blockDim.x) {
            const uint32_t c = b % C, q = b / C;
            const uint32_t m1 = ((q >> sl) << (sl + 1)) | (q & (h - 1));
            const uint32_t m2 = m1 + h;
            const size_t j = (size_t(q & (h - 1)) << a) + lo0 + c;
            const Fr w = ld8(tw + (j << (NTT_MAX_LOG - 1 - s)) * 8);
            const uint32_t e1 = m1 * C + c, e2 = m2 * C + c;
            const Fr x = lds_ld(lds, e1), y = lds_ld(lds, e2);
            lds_st(lds, e1, reduce(norm(add_w(x, y))));
            lds_st(lds, e2, mul(sub_w(x, y), w));
        }
        __syncthreads();
    }

    for (uint32_t e = threadIdx.x; e < tile_elems; e += blockDim.x) {
        const uint32_t mid = e / C, c = e % C;
        st8(out + (base + (size_t(mid) << a) + c) * 8, lds_ld(lds, e));
    }
}

template <bool CANON_IN>
__global__ void __launch_bounds__(PASS_THREADS) k_dif_pass(const uint32_t* __restrict__ in,
                                                      uint32_t* __restrict__ out, int L, int a, int S,
                                                      int C, const uint32_t* __restrict__ tw) {
    __shared__ uint32_t lds[LIMBS * TILE];
    dif_tile<CANON_IN>(lds, in, out, L, a, S, C, tw, blockIdx.x);
}

// out[row][k] = x[row][rev_L(k)] * scale, scale given as raw limbs: a plain
// integer leaves Montgomery form, a Montgomery value keeps it.
template <bool CANON_OUT>
__global__ void __launch_bounds__(THREADS) k_bitrev_scale(const uint32_t* __restrict__ in,
                                                          uint32_t* __restrict__ out, int L,
                                                          size_t total, Fr scale) {
    const size_t i = size_t(blockIdx.x) * THREADS + threadIdx.x;
    if (i >= total) return;
    const size_t n = size_t(1) << L;
    const size_t row = i >> L, k = i & (n - 1);
    const uint32_t src = L ? (__brev(uint32_t(k)) >> (32 - L)) : 0;
    Fr x = mul(ld8(in + (row * n + src) * 8), scale);
    if (CANON_OUT) x = canon(x);
    st8(out + i * 8, x);
}

__global__ void k_twiddles(const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                           uint32_t* __restrict__ tw, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    st8(tw + size_t(k) * 8, canon(mul(ld8(lo + (k & 1023) * 8), ld8(hi + (k >> 10) * 8))));
}

// ---- batches of transforms of several lengths -------------------------------

// Entry `e` of a fused batch, with its pass schedule (the same schedule as a
// single call of that length: pass p covers bits [a_p, a_p + S_p)).
struct NttBatchDesc {
    const uint32_t* in;
    uint32_t* out;
    uint32_t* work;
    uint32_t L, rows, passes, inverse;
    uint32_t a[NTT_BATCH_MAX_PASSES], S[NTT_BATCH_MAX_PASSES], C[NTT_BATCH_MAX_PASSES];
    uint32_t block_base[NTT_BATCH_MAX_PASSES];  // first block of this entry in pass p
    uint32_t elem_base;                          // first element in the bit-reversal launch
    uint32_t scale[8];                           // raw limbs of the output scale
};

__device__ __forceinline__ uint32_t find_entry(const NttBatchDesc* __restrict__ d, uint32_t m, int p, uint32_t key) {
    uint32_t lo = 0, hi = m - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        const uint32_t b = p < 0 ? d[mid].elem_base : d[mid].block_base[p];
        if (b <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <bool CANON_IN>
__global__ void __launch_bounds__(PASS_THREADS) k_dif_pass_batch(const NttBatchDesc* __restrict__ descs, uint32_t m,
                                                            int p, const uint32_t* __restrict__ tw_fwd,
                                                            const uint32_t* __restrict__ tw_inv) {
    __shared__ uint32_t lds[LIMBS * TILE];
    __shared__ uint32_t entry;
    if (threadIdx.x == 0) entry = find_entry(descs, m, p, blockIdx.x);
    __syncthreads();
    const NttBatchDesc& d = descs[entry];
    dif_tile<CANON_IN>(lds, CANON_IN ? d.in : d.work, d.work, int(d.L), int(d.a[p]), int(d.S[p]), int(d.C[p]),
                       d.inverse ? tw_inv : tw_fwd, blockIdx.x - d.block_base[p]);
}

__global__ void __launch_bounds__(THREADS) k_bitrev_batch(const NttBatchDesc* __restrict__ descs, uint32_t m,
                                                          uint32_t total) {
    const uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= total) return;
    const NttBatchDesc& d = descs[find_entry(descs, m, -1, i)];
    const uint32_t local = i - d.elem_base;
    const uint32_t L = d.L;
    const size_t n = size_t(1) << L;
    const size_t row = local >> L, k = local & (n - 1);
    const uint32_t src = L ? (__brev(uint32_t(k)) >> (32 - L)) : 0;
    const uint32_t* from = d.passes ? d.work : d.in;
    const Fr scale = unpack<FrParams>(d.scale);
    st8(d.out + size_t(local) * 8, canon(mul(ld8(from + (row * n + src) * 8), scale)));
}

Fr fr_from_words(const uint32_t (&w)[8]) { return to_mont<FrParams>(w); }

bool fr_is_one(const Fr& x) {
    const Fr c = canon(x), one = canon(fp_one<FrParams>());
    for (int i = 0; i < LIMBS; ++i)
        if (c.l[i] != one.l[i]) return false;
    return true;
}

}  // namespace

bool ntt_root_has_order_2_20(const uint32_t (&root)[8]) {
    if (!is_canonical<FrParams>(root)) return false;
    Fr p = fr_from_words(root);
    for (int i = 0; i < 19; ++i) p = sqr(p);
    if (fr_is_one(p)) return false;
    return fr_is_one(sqr(p));
}

void ntt_domain_prepare(NttDomain& d, const uint32_t (&root)[8], hipStream_t stream) {
    if (d.ready && std::memcmp(d.root_words, root, 32) == 0) return;
    require(ntt_root_has_order_2_20(root), "NTT root must have order 2^20");
    d.ready = false;
    const uint32_t count = 1u << (NTT_MAX_LOG - 1);
    const Fr w = fr_from_words(root);
    const Fr wi = inv(w);
    DeviceBuffer stage;
    std::vector<uint32_t> host(size_t(1024 + 512) * 8 * 2);
    for (int dir = 0; dir < 2; ++dir) {
        const Fr base = dir ? wi : w;
        uint32_t* lo = host.data() + size_t(dir) * (1536 * 8);
        uint32_t* hi = lo + 1024 * 8;
        Fr acc = fp_one<FrParams>();
        for (int j = 0; j < 1024; ++j) {
            uint32_t o[8];
            pack(canon(acc), o);
            std::memcpy(lo + j * 8, o, 32);
            acc = mul(acc, base);
        }
        const Fr step = acc;  // base^1024
        acc = fp_one<FrParams>();
        for (int j = 0; j < 512; ++j) {
            uint32_t o[8];
            pack(canon(acc), o);
            std::memcpy(hi + j * 8, o, 32);
            acc = mul(acc, step);
        }
    }
    uint32_t* dev = stage.as<uint32_t>(host.size());
    HIP_CHECK(hipMemcpyAsync(dev, host.data(), host.size() * 4, hipMemcpyHostToDevice, stream));
    uint32_t* f = d.fwd.as<uint32_t>(size_t(count) * 8);
    uint32_t* b = d.inv.as<uint32_t>(size_t(count) * 8);
    k_twiddles<<<grid_for(count, 256), 256, 0, stream>>>(dev, dev + 1024 * 8, f, count);
    HIP_CHECK_LAUNCH();
    k_twiddles<<<grid_for(count, 256), 256, 0, stream>>>(dev + 1536 * 8, dev + 1536 * 8 + 1024 * 8, b, count);
    HIP_CHECK_LAUNCH();
    HIP_CHECK(hipStreamSynchronize(stream));  // `stage` and `host` die here
    std::memcpy(d.root_words, root, 32);
    d.ready = true;
}

// Scale applied by the bit-reversal pass to Montgomery work values:
// canonical output wants a plain integer (1 or n^-1); Montgomery output wants
// the Montgomery form of the same value.
static Fr output_scale(int L, bool inverse, bool canonical) {
    Fr scale = fp_zero<FrParams>();
    if (canonical) {
        if (!inverse) {
            scale.l[0] = 1;
        } else {
            uint32_t nw[8] = {0};
            nw[L / 32] = 1u << (L % 32);
            const Fr ninv_m = inv(to_mont<FrParams>(nw));
            uint32_t plain[8];
            from_mont(ninv_m, plain);
            scale = unpack<FrParams>(plain);
        }
    } else {
        if (!inverse) {
            scale = fp_one<FrParams>();
        } else {
            uint32_t nw[8] = {0};
            nw[L / 32] = 1u << (L % 32);
            scale = canon(inv(to_mont<FrParams>(nw)));
        }
    }
    return scale;
}

static void run_rows(const NttDomain& d, const uint32_t* in, uint32_t* out, int L, size_t rows,
                     bool inverse, bool canonical, NttWorkspace& ws, hipStream_t stream) {
    require(d.ready, "NTT domain not prepared");
    require(L >= 0 && L <= NTT_MAX_LOG, "NTT length must be a power of two <= 2^20");
    const size_t n = size_t(1) << L;
    const size_t total = n * rows;
    if (!total) return;
    uint32_t* work = ws.a.as<uint32_t>(total * 8);
    const uint32_t* tw = static_cast<const uint32_t*>(inverse ? d.inv.ptr : d.fwd.ptr);
    if (L == 0) {
        HIP_CHECK(hipMemcpyAsync(out, in, total * 32, hipMemcpyDeviceToDevice, stream));
        return;
    }
    const int max_stages = d.pass_stages[L] >= 1 && d.pass_stages[L] <= MAX_STAGES ? d.pass_stages[L] : MAX_STAGES;
    const int passes = (L + max_stages - 1) / max_stages;
    int top = L;
    bool first = true;
    for (int p = 0; p < passes; ++p) {
        const int remaining_passes = passes - p;
        const int S = (top + remaining_passes - 1) / remaining_passes;
        const int a = top - S;
        const int C = int(std::min<size_t>(size_t(1) << a, size_t(TILE >> S)));
        const size_t blocks = rows * (n / (size_t(1u << S) * C));
        if (first && canonical)
            k_dif_pass<true><<<unsigned(blocks), pass_threads((1u << S) * C), 0, stream>>>(in, work, L, a, S, C, tw);
        else
            k_dif_pass<false><<<unsigned(blocks), pass_threads((1u << S) * C), 0, stream>>>(first ? in : work, work, L,
                                                                                           a, S, C, tw);
        HIP_CHECK_LAUNCH();
        first = false;
        top = a;
    }
    const Fr scale = output_scale(L, inverse, canonical);
    if (canonical)
        k_bitrev_scale<true><<<grid_for(total, THREADS), THREADS, 0, stream>>>(work, out, L, total, scale);
    else
        k_bitrev_scale<false><<<grid_for(total, THREADS), THREADS, 0, stream>>>(work, out, L, total, scale);
    HIP_CHECK_LAUNCH();
}

void ntt_batch_prepare(const NttDomain& d, const std::vector<NttBatchEntry>& entries, NttBatchPlan& plan,
                       hipStream_t stream) {
    require(d.ready, "NTT domain not prepared");
    require(!entries.empty(), "empty NTT batch");
    // Outputs may not overlap any other entry's input or output (the entries
    // run together); an entry may transform in place.  Checked in
    // O(E log E): the output ranges sorted by start must be disjoint, and
    // every input range may only meet its own entry's output.
    struct Range {
        uintptr_t lo, hi;
        size_t entry;
    };
    std::vector<Range> outs;
    outs.reserve(entries.size());
    for (size_t i = 0; i < entries.size(); ++i) {
        const NttBatchEntry& x = entries[i];
        require(x.log_n >= 0 && x.log_n <= NTT_MAX_LOG && x.rows >= 1 && x.in && x.out, "invalid NTT batch entry");
        const uintptr_t o = reinterpret_cast<uintptr_t>(x.out);
        outs.push_back({o, o + (x.rows << x.log_n) * 32, i});
    }
    std::sort(outs.begin(), outs.end(), [](const Range& a, const Range& b) { return a.lo < b.lo; });
    for (size_t k = 1; k < outs.size(); ++k) require(outs[k - 1].hi <= outs[k].lo, "NTT batch entries overlap");
    for (size_t j = 0; j < entries.size(); ++j) {
        const uintptr_t lo = reinterpret_cast<uintptr_t>(entries[j].in),
                        hi = lo + (entries[j].rows << entries[j].log_n) * 32;
        // first output range that ends after lo (ends are sorted as the ranges are disjoint)
        size_t a = 0, b = outs.size();
        while (a < b) {
            const size_t m = (a + b) / 2;
            if (outs[m].hi <= lo) a = m + 1;
            else b = m;
        }
        for (size_t k = a; k < outs.size() && outs[k].lo < hi; ++k)
            require(outs[k].entry == j, "NTT batch entries overlap");
    }
    std::vector<NttBatchDesc> descs(entries.size());
    Fr scales[2][NTT_MAX_LOG + 1];
    bool have_scale[2][NTT_MAX_LOG + 1] = {};
    uint64_t work = 0, elems = 0;
    uint32_t blocks[NTT_BATCH_MAX_PASSES] = {0}, tile_max[NTT_BATCH_MAX_PASSES] = {0};
    int max_passes = 0;
    for (size_t i = 0; i < entries.size(); ++i) {
        const NttBatchEntry& x = entries[i];
        NttBatchDesc& e = descs[i];
        const int L = x.log_n;
        const size_t n = size_t(1) << L;
        int stages = d.pass_stages[L] >= 1 && d.pass_stages[L] <= MAX_STAGES ? d.pass_stages[L] : MAX_STAGES;
        stages = std::max(stages, (L + NTT_BATCH_MAX_PASSES - 1) / NTT_BATCH_MAX_PASSES);
        const int passes = L ? (L + stages - 1) / stages : 0;
        e.in = x.in;
        e.out = x.out;
        e.L = uint32_t(L);
        e.rows = uint32_t(x.rows);
        e.passes = uint32_t(passes);
        e.inverse = x.inverse ? 1u : 0u;
        e.work = reinterpret_cast<uint32_t*>(work);  // offset until the workspace is known
        int top = L;
        for (int p = 0; p < NTT_BATCH_MAX_PASSES; ++p) {
            e.block_base[p] = blocks[p];
            if (p >= passes) continue;
            const int remaining = passes - p;
            const int S = (top + remaining - 1) / remaining;
            const int a = top - S;
            const int C = int(std::min<size_t>(size_t(1) << a, size_t(TILE >> S)));
            e.a[p] = uint32_t(a);
            e.S[p] = uint32_t(S);
            e.C[p] = uint32_t(C);
            const uint64_t nb = uint64_t(x.rows) * (n / (size_t(1u << S) * C));
            tile_max[p] = std::max<uint32_t>(tile_max[p], (1u << S) * uint32_t(C));
            require(uint64_t(blocks[p]) + nb < (uint64_t(1) << 31), "NTT batch too large");
            blocks[p] += uint32_t(nb);
            top = a;
        }
        max_passes = std::max(max_passes, passes);
        e.elem_base = uint32_t(elems);
        elems += uint64_t(x.rows) << L;
        require(elems < (uint64_t(1) << 31), "NTT batch too large");
        if (passes) work += uint64_t(x.rows) << L;
        // length 1: identity; otherwise 1 or n^-1, one inversion per distinct length
        if (passes && !have_scale[x.inverse][L]) {
            scales[x.inverse][L] = output_scale(L, x.inverse, true);
            have_scale[x.inverse][L] = true;
        }
        pack(passes ? scales[x.inverse][L] : fp_one<FrParams>(), e.scale);
    }
    // Growing a device buffer frees one that queued work may still use.
    if (std::max<uint64_t>(work, 1) * 32 > plan.work.bytes || descs.size() * sizeof(NttBatchDesc) > plan.descs.bytes)
        HIP_CHECK(hipStreamSynchronize(stream));
    uint32_t* wp = plan.work.as<uint32_t>(std::max<uint64_t>(work, 1) * 8);
    for (auto& e : descs) e.work = wp + reinterpret_cast<uintptr_t>(e.work) * 8;
    plan.count = uint32_t(entries.size());
    plan.max_passes = max_passes;
    plan.total_elems = uint32_t(elems);
    for (int p = 0; p < NTT_BATCH_MAX_PASSES; ++p) {
        plan.pass_blocks[p] = blocks[p];
        plan.pass_threads[p] = pass_threads(tile_max[p]);
    }
    // The previous upload from the pinned descriptors may still be pending.
    if (plan.upload_pending) {
        HIP_CHECK(hipEventSynchronize(plan.uploaded));
        plan.upload_pending = false;
    }
    auto* host = plan.host_descs.as<NttBatchDesc>(descs.size());
    std::memcpy(host, descs.data(), descs.size() * sizeof(NttBatchDesc));
    auto* dev = plan.descs.as<NttBatchDesc>(descs.size());
    HIP_CHECK(hipMemcpyAsync(dev, host, descs.size() * sizeof(NttBatchDesc), hipMemcpyHostToDevice, stream));
    if (!plan.uploaded) HIP_CHECK(hipEventCreateWithFlags(&plan.uploaded, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(plan.uploaded, stream));
    plan.upload_pending = true;
}

void ntt_batch_enqueue(const NttDomain& d, const NttBatchPlan& plan, hipStream_t stream) {
    require(d.ready && plan.count, "NTT batch not prepared");
    const auto* descs = static_cast<const NttBatchDesc*>(plan.descs.ptr);
    const auto* fwd = static_cast<const uint32_t*>(d.fwd.ptr);
    const auto* inv_tw = static_cast<const uint32_t*>(d.inv.ptr);
    for (int p = 0; p < plan.max_passes; ++p) {
        if (!plan.pass_blocks[p]) continue;
        if (p == 0)
            k_dif_pass_batch<true><<<plan.pass_blocks[p], plan.pass_threads[p], 0, stream>>>(descs, plan.count, p, fwd,
                                                                                              inv_tw);
        else
            k_dif_pass_batch<false><<<plan.pass_blocks[p], plan.pass_threads[p], 0, stream>>>(descs, plan.count, p,
                                                                                               fwd, inv_tw);
        HIP_CHECK_LAUNCH();
    }
    k_bitrev_batch<<<grid_for(plan.total_elems, THREADS), THREADS, 0, stream>>>(descs, plan.count, plan.total_elems);
    HIP_CHECK_LAUNCH();
}

void ntt_rows(const NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
              bool inverse, NttWorkspace& ws, hipStream_t stream) {
    run_rows(d, in, out, log_n, rows, inverse, true, ws, stream);
}

void ntt_rows_mont(const NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
                   bool inverse, NttWorkspace& ws, hipStream_t stream) {
    run_rows(d, in, out, log_n, rows, inverse, false, ws, stream);
}

}  // namespace dvgpu
