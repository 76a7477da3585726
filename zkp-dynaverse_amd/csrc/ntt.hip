// Natural-order radix-2 NTT over BN254 Fr (decimation in frequency).
//
// A transform of length n = 2^L runs as one LDS pass (L <= 11) or as two
// (the four-step split n = 2^S1 x 2^a, S1 = L/2, a = L - S1):
//   pass 1: for every column x2 < 2^a, the S1 DIF stages over the high index
//           bits with the full twiddles omega_(2^(a+s+1))^(j 2^a + x2); the
//           column's outputs leave in natural order, row k1 = its frequency;
//   pass 2: for every row k1, the a DIF stages over the row with the
//           sub-transform twiddles; output k2 of row k1 is X[k1 + 2^S1 k2].
// Inside a pass every sub-transform sits in LDS (nine 29-bit limb planes), its
// DIF leaves frequency k at position bitrev(k), and the store reads position
// bitrev(k) for output k: no separate permutation pass exists.  A tile holds
// U sub-transforms side by side (U columns, or U rows), so pass 1 reads and
// writes runs of U elements and pass 2 writes runs of U elements.
//
// The transform is linear and its twiddles are Montgomery values, so it maps
// the raw words of its input to the raw words of its output whatever form
// they are in: canonical in gives canonical out, Montgomery form in gives
// Montgomery form out, with no conversion product.  The inverse applies 1/n
// with one product per element as the last pass stores.
//
// Butterflies: (x, y) -> (x + y, (x - y) w), x - y formed limb-wise without a
// carry pass and fed straight into the Montgomery product.  Launches with
// enough work run two stages per LDS round trip (radix-4 groups of four
// elements in registers, two lazy sums reduced once); small launches keep
// one stage per round trip, which doubles the threads at work.  The stage
// tables hold twiddles pre-split into limbs; the last round of a
// sub-transform with twiddles of order 2 and 4 skips its unit products.
#include "ntt.hpp"
#include "ntt_tile.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

namespace dvgpu {

using namespace bn254;

namespace {

using namespace ntt;

constexpr int THREADS = 256;  // element-wise kernels

template <bool R4>
__global__ void __launch_bounds__(R4 ? 512 : 1024) k_ntt_pass(NttPass p) {
    __shared__ uint32_t lds[LIMBS * TILE];
    ntt_tile<R4>(lds, p, xcd_order(blockIdx.x, gridDim.x));
}

// ---- batches of transforms of several lengths -------------------------------

struct NttBatchDesc {
    NttPass pass[MAX_PASSES];
    uint32_t block_base[MAX_PASSES];  // first block of this entry in pass p
};

// Last entry whose first block of pass p is at or below `key` (an entry with
// no block in pass p shares its base with the next one, which wins).
__device__ __forceinline__ uint32_t find_entry(const NttBatchDesc* __restrict__ d, uint32_t m, int p, uint32_t key) {
    uint32_t lo = 0, hi = m - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (d[mid].block_base[p] <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <bool R4>
__global__ void __launch_bounds__(R4 ? 512 : 1024) k_ntt_pass_batch(const NttBatchDesc* __restrict__ descs,
                                                                    uint32_t m, int p) {
    __shared__ uint32_t lds[LIMBS * TILE];
    __shared__ uint32_t entry;
    const uint32_t lb = xcd_order(blockIdx.x, gridDim.x);
    if (threadIdx.x == 0) entry = find_entry(descs, m, p, lb);
    __syncthreads();
    const NttBatchDesc& d = descs[entry];
    ntt_tile<R4>(lds, d.pass[p], lb - d.block_base[p]);
}

// ---- twiddle tables ------------------------------------------------------------

__global__ void k_twiddles(const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                           uint32_t* __restrict__ tw, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    st8(tw + size_t(k) * 8, canon(mul(ld8(lo + (k & 1023) * 8), ld8(hi + (k >> 10) * 8))));
}

__global__ void k_stage_table(const uint32_t* __restrict__ pw, uint32_t* __restrict__ out, uint32_t total) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < total) stage_table_entry(pw, out, e);
}

__global__ void k_pass1_table(const uint32_t* __restrict__ pw, uint32_t* __restrict__ out, uint32_t L, uint32_t S,
                              uint32_t logCT, uint32_t total) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < total) pass1_table_entry(pw, out, L, S, logCT, e);
}

Fr fr_from_words(const uint32_t (&w)[8]) { return to_mont<FrParams>(w); }

bool fr_is_one(const Fr& x) {
    const Fr c = canon(x), one = canon(fp_one<FrParams>());
    for (int i = 0; i < LIMBS; ++i)
        if (c.l[i] != one.l[i]) return false;
    return true;
}

// The Montgomery limbs of 1/2^L, from a table made once per process (a
// Fermat inversion on the host per inverse transform would hold the stream
// for several microseconds before every launch).
Fr inverse_scale(int L) {
    static const std::array<Fr, NTT_MAX_LOG + 1> table = [] {
        std::array<Fr, NTT_MAX_LOG + 1> t{};
        for (int l = 0; l <= NTT_MAX_LOG; ++l) {
            uint32_t nw[8] = {0};
            nw[l / 32] = 1u << (l % 32);
            t[size_t(l)] = canon(inv(to_mont<FrParams>(nw)));
        }
        return t;
    }();
    return table[size_t(L)];
}

// The pass-1 table of length 2^L for one direction, built on first use.
const uint32_t* pass1_table(NttTables& t, int L, hipStream_t stream) {
    DeviceBuffer& b = t.pass1[L];
    if (!b.ptr) {
        const uint32_t total = pass1_table_entries(L);
        uint32_t* out = b.as<uint32_t>(size_t(total) * LIMBS);
        k_pass1_table<<<grid_for(total, THREADS), THREADS, 0, stream>>>(static_cast<const uint32_t*>(t.pow.ptr), out,
                                                                         uint32_t(L), uint32_t(L / 2),
                                                                         pass1_log_ct(L), total);
        HIP_CHECK_LAUNCH();
    }
    return static_cast<const uint32_t*>(b.ptr);
}

// The passes of one entry, tables built as needed.
LengthPlan plan_entry(NttDomain& d, const uint32_t* in, uint32_t* out, uint32_t* work, int L, uint64_t rows,
                      bool inverse, hipStream_t stream) {
    require(L >= 0 && L <= NTT_MAX_LOG, "NTT length must be a power of two <= 2^20");
    require(rows >= 1 && rows < (uint64_t(1) << 31), "NTT row count out of range");
    NttTables& t = inverse ? d.inv : d.fwd;
    const Fr scale = inverse && L ? inverse_scale(L) : fp_one<FrParams>();
    const uint32_t* p1 = L > SINGLE_MAX_LOG ? pass1_table(t, L, stream) : nullptr;
    const LengthPlan lp = plan_length(in, out, work, L, rows, static_cast<const uint32_t*>(t.stage.ptr), p1,
                                      inverse && L ? &scale : nullptr);
    for (int p = 0; p < lp.passes; ++p)
        require(lp.pass[p].mode == PASS_ROWS || rows * lp.pass[p].tiles < (uint64_t(1) << 31), "NTT batch too large");
    return lp;
}

// Stages per LDS round trip for a launch: the tuning table's choice for the
// length (1: radix-2 rounds, 2 or more: radix-4 rounds), else by size.
bool radix4_for(const NttDomain& d, int L, uint64_t blocks, uint32_t tile) {
    const int forced = d.pass_stages[L];
    return forced ? forced >= 2 : use_radix4(blocks, tile);
}

void launch_pass(const NttPass& p, bool r4, uint32_t blocks, uint32_t tile, hipStream_t stream) {
    if (r4)
        k_ntt_pass<true><<<blocks, threads_for(tile, true), 0, stream>>>(p);
    else
        k_ntt_pass<false><<<blocks, threads_for(tile, false), 0, stream>>>(p);
    HIP_CHECK_LAUNCH();
}

}  // namespace

bool ntt_root_has_order_2_20(const uint32_t (&root)[8]) {
    if (!is_canonical<FrParams>(root)) return false;
    Fr p = fr_from_words(root);
    for (int i = 0; i < 19; ++i) p = sqr(p);
    if (fr_is_one(p)) return false;
    return fr_is_one(sqr(p));
}

uint64_t ntt_domain_bytes(int max_log) {
    uint64_t one = DeviceBuffer::grown(size_t(1u << (NTT_MAX_LOG - 1)) * 32) +
                   DeviceBuffer::grown(size_t((1u << NTT_STAGE_TABLE_LOG) - 1) * LIMBS * 4);
    for (int L = SINGLE_MAX_LOG + 1; L <= std::min(max_log, NTT_MAX_LOG); ++L)
        one += DeviceBuffer::grown(size_t(pass1_table_entries(L)) * LIMBS * 4);
    return 2 * one;
}

void ntt_domain_prepare(NttDomain& d, const uint32_t (&root)[8], hipStream_t stream) {
    if (d.ready && std::memcmp(d.root_words, root, 32) == 0) return;
    require(ntt_root_has_order_2_20(root), "NTT root must have order 2^20");
    d.ready = false;
    for (auto* t : {&d.fwd, &d.inv})
        for (auto& b : t->pass1) b.release();  // tables of another root
    const uint32_t count = 1u << (NTT_MAX_LOG - 1);
    const Fr w = fr_from_words(root);
    const Fr wi = inv(w);
    DeviceBuffer staging;
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
    uint32_t* dev = staging.as<uint32_t>(host.size());
    HIP_CHECK(hipMemcpyAsync(dev, host.data(), host.size() * 4, hipMemcpyHostToDevice, stream));
    const uint32_t stage_total = (1u << NTT_STAGE_TABLE_LOG) - 1;
    for (int dir = 0; dir < 2; ++dir) {
        NttTables& t = dir ? d.inv : d.fwd;
        uint32_t* pw = t.pow.as<uint32_t>(size_t(count) * 8);
        const uint32_t* src = dev + size_t(dir) * 1536 * 8;
        k_twiddles<<<grid_for(count, THREADS), THREADS, 0, stream>>>(src, src + 1024 * 8, pw, count);
        HIP_CHECK_LAUNCH();
        uint32_t* st = t.stage.as<uint32_t>(size_t(stage_total) * LIMBS);
        k_stage_table<<<grid_for(stage_total, THREADS), THREADS, 0, stream>>>(pw, st, stage_total);
        HIP_CHECK_LAUNCH();
    }
    HIP_CHECK(hipStreamSynchronize(stream));  // `staging` and `host` die here
    std::memcpy(d.root_words, root, 32);
    d.ready = true;
}

static void run_rows(NttDomain& d, const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse,
                     NttWorkspace& ws, hipStream_t stream) {
    require(d.ready, "NTT domain not prepared");
    require(L >= 0 && L <= NTT_MAX_LOG, "NTT length must be a power of two <= 2^20");
    if (!rows) return;
    uint32_t* work = L > SINGLE_MAX_LOG ? ws.a.as<uint32_t>((rows << L) * 8) : nullptr;
    const LengthPlan lp = plan_entry(d, in, out, work, L, rows, inverse, stream);
    for (int p = 0; p < lp.passes; ++p)
        launch_pass(lp.pass[p], radix4_for(d, L, lp.blocks[p], lp.tile[p]), lp.blocks[p], lp.tile[p], stream);
}

Fr ntt_inverse_scale(int log_n) {
    require(log_n >= 0 && log_n <= NTT_MAX_LOG, "NTT length must be a power of two <= 2^20");
    return inverse_scale(log_n);
}

void ntt_rows(NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows, bool inverse,
              NttWorkspace& ws, hipStream_t stream) {
    run_rows(d, in, out, log_n, rows, inverse, ws, stream);
}

void ntt_rows_mont(NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows, bool inverse,
                   NttWorkspace& ws, hipStream_t stream) {
    run_rows(d, in, out, log_n, rows, inverse, ws, stream);
}

void ntt_batch_prepare(NttDomain& d, const std::vector<NttBatchEntry>& entries, NttBatchPlan& plan,
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
    // Work area: every two-pass entry gets its own slice (offsets first,
    // pointers once the buffer is sized).
    uint64_t work = 0;
    std::vector<uint64_t> work_at(entries.size(), 0);
    for (size_t i = 0; i < entries.size(); ++i)
        if (entries[i].log_n > SINGLE_MAX_LOG) {
            work_at[i] = work;
            work += uint64_t(entries[i].rows) << entries[i].log_n;
        }
    // Growing a device buffer frees one that queued work may still use.
    const size_t desc_bytes = entries.size() * sizeof(NttBatchDesc);
    if (std::max<uint64_t>(work, 1) * 32 > plan.work.bytes || desc_bytes > plan.descs.bytes)
        HIP_CHECK(hipStreamSynchronize(stream));
    uint32_t* wp = plan.work.as<uint32_t>(std::max<uint64_t>(work, 1) * 8);
    std::vector<NttBatchDesc> descs(entries.size());
    uint64_t blocks[NTT_BATCH_MAX_PASSES] = {0, 0}, groups[NTT_BATCH_MAX_PASSES] = {0, 0};
    uint32_t tile_max[NTT_BATCH_MAX_PASSES] = {0, 0};
    int forced[NTT_BATCH_MAX_PASSES] = {-1, -1};  // the entries' common tuned choice, -2 when they differ
    int max_passes = 0;
    for (size_t i = 0; i < entries.size(); ++i) {
        const NttBatchEntry& x = entries[i];
        const LengthPlan lp = plan_entry(d, x.in, x.out, wp + work_at[i] * 8, x.log_n, x.rows, x.inverse, stream);
        NttBatchDesc& e = descs[i];
        for (int p = 0; p < NTT_BATCH_MAX_PASSES; ++p) {
            e.block_base[p] = uint32_t(blocks[p]);
            if (p >= lp.passes) {
                e.pass[p] = lp.pass[0];  // never read: this entry has no block in pass p
                continue;
            }
            e.pass[p] = lp.pass[p];
            const int f = d.pass_stages[x.log_n] ? (d.pass_stages[x.log_n] >= 2 ? 2 : 1) : 0;
            forced[p] = forced[p] == -1 || forced[p] == f ? f : -2;
            blocks[p] += lp.blocks[p];
            groups[p] += uint64_t(lp.blocks[p]) * (lp.tile[p] >> 2);
            tile_max[p] = std::max(tile_max[p], lp.tile[p]);
            require(blocks[p] < (uint64_t(1) << 31), "NTT batch too large");
        }
        max_passes = std::max(max_passes, lp.passes);
    }
    plan.count = uint32_t(entries.size());
    plan.max_passes = max_passes;
    for (int p = 0; p < NTT_BATCH_MAX_PASSES; ++p) {
        plan.pass_blocks[p] = uint32_t(blocks[p]);
        plan.pass_radix4[p] = forced[p] > 0 ? forced[p] == 2 : groups[p] >= R4_MIN_GROUPS;
        plan.pass_threads[p] = threads_for(std::max<uint32_t>(tile_max[p], 2), plan.pass_radix4[p]);
    }
    // The previous upload from the pinned descriptors may still be pending.
    if (plan.upload_pending) {
        HIP_CHECK(hipEventSynchronize(plan.uploaded));
        plan.upload_pending = false;
    }
    auto* host = plan.host_descs.as<NttBatchDesc>(descs.size());
    std::memcpy(host, descs.data(), desc_bytes);
    auto* dev = plan.descs.as<NttBatchDesc>(descs.size());
    HIP_CHECK(hipMemcpyAsync(dev, host, desc_bytes, hipMemcpyHostToDevice, stream));
    if (!plan.uploaded) create_event(plan.uploaded, hipEventDisableTiming);
    HIP_CHECK(hipEventRecord(plan.uploaded, stream));
    plan.upload_pending = true;
}

void ntt_batch_enqueue(const NttDomain& d, const NttBatchPlan& plan, hipStream_t stream) {
    require(d.ready && plan.count, "NTT batch not prepared");
    const auto* descs = static_cast<const NttBatchDesc*>(plan.descs.ptr);
    for (int p = 0; p < plan.max_passes; ++p) {
        if (!plan.pass_blocks[p]) continue;
        if (plan.pass_radix4[p])
            k_ntt_pass_batch<true><<<plan.pass_blocks[p], plan.pass_threads[p], 0, stream>>>(descs, plan.count, p);
        else
            k_ntt_pass_batch<false><<<plan.pass_blocks[p], plan.pass_threads[p], 0, stream>>>(descs, plan.count, p);
        HIP_CHECK_LAUNCH();
    }
}

}  // namespace dvgpu
