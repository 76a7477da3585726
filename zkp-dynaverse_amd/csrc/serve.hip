// Serving form: one call applies a batch of sparse updates to device-resident
// session states and regenerates each session's commitment, opening and
// evaluations.
//
// A session of size n = 2^L keeps, in two caller-owned device arenas:
//   points: window tables of its basis P_i and opening basis U_i, row i
//           holding [2^(c w)] B_i for w < ceil(255 / c) (Montgomery affine),
//           and two rows holding the cached commitment C and opening W as
//           XYZZ points (128 bytes each);
//   fr:     coefficients f (canonical), evaluations NTT(f), the linear-term
//           vector l (the opening functional, l_i = z^i) and the cached value
//           y = <l, f>.
// An update {(i_j, d_j)} becomes, for every session of the batch at once:
//   C <- C + sum_j d_j P_{i_j},  W <- W + sum_j d_j U_{i_j}
//   f_{i_j} += d_j,  y += sum_j d_j l_{i_j},  evaluations <- NTT(f).
// With the window tables every digit of every window lands in one bucket set
// (no doublings between windows), and the finishing kernel adds the bucket
// sum to the cached XYZZ point in place.  The host turns the new C and W into
// canonical affine points with one batched inversion.
#include <algorithm>
#include <cstring>
#include <iterator>
#include <vector>

#include "../../include/dynaverse_hip_ext.h"
#include "owner.hpp"

namespace dvgpu {

using namespace bn254;

namespace {

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

// Per update, as the kernels read it.
struct UpdDesc {
    uint32_t entry_base;   // first changed entry of this update in the batch
    uint32_t k;
    uint32_t scalar_base;  // first delta in the batch's scalar array
    uint32_t pad;
    uint64_t coeff_offset, linear_offset, value_slot;
};

__device__ __forceinline__ uint32_t find_update(const UpdDesc* __restrict__ d, uint32_t m, uint32_t e) {
    uint32_t lo = 0, hi = m - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (d[mid].entry_base <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ Fr ld_fr(const uint32_t* p) {
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return unpack<FrParams>(w);
}

__device__ __forceinline__ void st_fr(uint32_t* p, const Fr& v) {
    uint32_t w[8];
    pack(canon(v), w);
    reinterpret_cast<uint4*>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// f[coeff_offset + i_j] += d_j (canonical in, canonical out); indices of one
// update are distinct and updates own disjoint coefficient ranges.
__global__ void k_scatter_add(const UpdDesc* __restrict__ descs, uint32_t m, uint32_t entries,
                              const uint32_t* __restrict__ local_index, const uint32_t* __restrict__ scalars,
                              uint32_t* __restrict__ fr) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= entries) return;
    const UpdDesc d = descs[find_update(descs, m, e)];
    const uint32_t j = e - d.entry_base;
    uint32_t* slot = fr + (d.coeff_offset + local_index[e]) * 8;
    st_fr(slot, add(ld_fr(slot), ld_fr(scalars + size_t(d.scalar_base + j) * 8)));
}

// One block per update: y += sum_j d_j l_{i_j}; the new y goes to the value
// slot and to out[r].
__global__ void __launch_bounds__(64) k_linear_dot(const UpdDesc* __restrict__ descs,
                                                   const uint32_t* __restrict__ local_index,
                                                   const uint32_t* __restrict__ scalars, uint32_t* __restrict__ fr,
                                                   uint32_t* __restrict__ out) {
    __shared__ uint32_t part[64 * 8];
    const UpdDesc d = descs[blockIdx.x];
    Fr acc = fp_zero<FrParams>();
    for (uint32_t j = threadIdx.x; j < d.k; j += 64) {
        uint32_t dw[8];
        const uint32_t* s = scalars + size_t(d.scalar_base + j) * 8;
        for (int t = 0; t < 8; ++t) dw[t] = s[t];
        const Fr l = ld_fr(fr + (d.linear_offset + local_index[d.entry_base + j]) * 8);
        acc = add(acc, mul(to_mont<FrParams>(dw), l));  // (d R) l R^-1 = d l
    }
    uint32_t w[8];
    pack(canon(acc), w);
    for (int t = 0; t < 8; ++t) part[threadIdx.x * 8 + t] = w[t];
    __syncthreads();
    if (threadIdx.x == 0) {
        Fr y = ld_fr(fr + d.value_slot * 8);
        for (int t = 0; t < 64; ++t) y = add(y, ld_fr(part + t * 8));
        st_fr(fr + d.value_slot * 8, y);
        st_fr(out + size_t(blockIdx.x) * 8, y);
    }
}

__device__ __forceinline__ Fq ld_fq(const uint32_t* p) {
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return unpack<FqParams>(w);
}

__device__ __forceinline__ void st_fq(uint32_t* p, const Fq& v) {
    uint32_t w[8];
    pack(v, w);
    reinterpret_cast<uint4*>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

__device__ __forceinline__ Xyzz<FqOps> ld_xyzz(const uint32_t* p) {
    return {ld_fq(p), ld_fq(p + 8), ld_fq(p + 16), ld_fq(p + 24)};
}

__device__ __forceinline__ void st_xyzz(uint32_t* p, const Xyzz<FqOps>& v) {
    st_fq(p, v.X);
    st_fq(p + 8, v.Y);
    st_fq(p + 16, v.ZZ);
    st_fq(p + 24, v.ZZZ);
}

// One lane per MSM (window-table mode, one window value each):
// cache <- cache + V, out <- cache, unless the batch was flagged.
__global__ void k_add_to_cache(const MsmDesc* __restrict__ descs, uint32_t m, const uint32_t* __restrict__ windows,
                               uint32_t* __restrict__ points, const uint32_t* __restrict__ gate,
                               uint32_t* __restrict__ out) {
    constexpr int XW = PointLayout<FqOps>::XYZZ_WORDS, AW = PointLayout<FqOps>::AFFINE_WORDS;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const MsmDesc d = descs[i];
    uint32_t* cache = points + size_t(d.result_slot) * AW;
    const Xyzz<FqOps> r = xyzz_add<FqOps>(ld_xyzz(cache), ld_xyzz(windows + size_t(d.window_base) * XW));
    if (gate[0] == 0) st_xyzz(cache, r);
    st_xyzz(out + size_t(d.out_index) * XW, r);
}

// XYZZ (Montgomery words) -> canonical affine bytes for a batch, with one
// field inversion (Montgomery's trick over the ZZZ coordinates): x = X ZZ^-1,
// y = Y ZZZ^-1 and ZZ^-1 = (ZZ ZZZ^-1)^2 since ZZ^3 = ZZZ^2.
void batch_to_affine(const uint32_t* xyzz, size_t count, uint8_t* out, size_t out_stride) {
    auto ld = [](const uint32_t* p) {
        uint32_t w[8];
        std::memcpy(w, p, 32);
        return unpack<FqParams>(w);
    };
    std::vector<Fq> zzz(count), pre(count);
    std::vector<bool> inf(count);
    Fq acc = fp_one<FqParams>();
    for (size_t i = 0; i < count; ++i) {
        const uint32_t* p = xyzz + i * 32;
        zzz[i] = ld(p + 24);
        inf[i] = is_zero(ld(p + 16)) || is_zero(zzz[i]);
        pre[i] = acc;
        if (!inf[i]) acc = mul(acc, zzz[i]);
    }
    Fq inv_acc = inv(acc);
    for (size_t i = count; i-- > 0;) {
        uint8_t* o = out + i * out_stride;
        if (inf[i]) {
            std::memset(o, 0, 64);
            continue;
        }
        const Fq zzz_inv = mul(inv_acc, pre[i]);
        inv_acc = mul(inv_acc, zzz[i]);
        const uint32_t* p = xyzz + i * 32;
        const Fq zz_inv = sqr(mul(ld(p + 16), zzz_inv));
        uint32_t xw[8], yw[8];
        from_mont(mul(ld(p), zz_inv), xw);
        from_mont(mul(ld(p + 8), zzz_inv), yw);
        std::memcpy(o, xw, 32);
        std::memcpy(o + 32, yw, 32);
    }
}

// Update MSMs are small and latency-bound (tens of points per bucket set of
// 128): short accumulation tasks and one bucket per reduction thread shorten
// the serial chains; k_heavy and the bucket sums absorb the extra partials.
MsmConfig update_msm_config(const Owner* o) {
    MsmConfig cfg = o->msm_config;
    cfg.task_points = 4;
    cfg.chunk = 1;
    return cfg;
}

// The update shape wins over per-size reduction chunks tuned for large MSMs.
WindowTable update_window_table(const Owner* o) {
    WindowTable t = o->window_table;
    for (int& ch : t.chunk) ch = 0;
    return t;
}

struct Interval {
    uint64_t lo, hi;  // [lo, hi)
};

// Sorts `v` by start; true if any two of its intervals intersect.
bool sort_and_overlap(std::vector<Interval>& v) {
    std::sort(v.begin(), v.end(), [](const Interval& a, const Interval& b) { return a.lo < b.lo; });
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i].lo < v[i - 1].hi) return true;
    return false;
}

// True if [lo, hi) meets one of the disjoint intervals of `sorted`.
bool hits(const std::vector<Interval>& sorted, uint64_t lo, uint64_t hi) {
    auto it = std::upper_bound(sorted.begin(), sorted.end(), lo,
                               [](uint64_t x, const Interval& s) { return x < s.lo; });
    if (it != sorted.begin() && std::prev(it)->hi > lo) return true;  // starts at or before lo
    return it != sorted.end() && it->lo < hi;                         // starts inside [lo, hi)
}

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

extern "C" {

int dvgpu_ext_update_batch(void* owner, void* points, size_t point_count, int window_bits, void* fr,
                           size_t fr_count, const DvgpuUpdate* updates, size_t count, const uint8_t* root,
                           uint8_t* results, float* device_ms, char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(points && fr && updates && count && root && results && point_count <= 0xffffffffull &&
                    count <= (1u << 20) && window_bits >= 4 && window_bits <= 20,
                "invalid update batch arguments");
        uint32_t rw[8];
        std::memcpy(rw, root, 32);
        require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
        const uint64_t W = uint64_t((255 + window_bits - 1) / window_bits);  // points per table row
        // ---- validation: ranges, distinct indices, no two updates sharing state
        std::vector<Interval> written_fr, read_fr, written_pts, read_pts;
        uint64_t entries = 0, scalars = 0, msm_points = 0;
        for (size_t r = 0; r < count; ++r) {
            const DvgpuUpdate& u = updates[r];
            require(u.log_n <= uint32_t(NTT_MAX_LOG), "session size must be at most 2^20");
            const uint64_t n = uint64_t(1) << u.log_n;
            require(u.k <= n && (!u.k || (u.indices && u.deltas)), "invalid update entries");
            require((uint64_t(u.srs_row) + n) * W <= point_count && (uint64_t(u.opening_row) + n) * W <= point_count &&
                        uint64_t(u.commitment_row) * W + 2 <= point_count &&
                        uint64_t(u.opening_cache_row) * W + 2 <= point_count &&
                        u.commitment_row != u.opening_cache_row,
                    "session points out of range");
            require(u.coeff_offset + n <= fr_count && u.eval_offset + n <= fr_count &&
                        u.linear_offset + n <= fr_count && u.value_slot < fr_count,
                    "session field elements out of range");
            std::vector<uint32_t> idx(u.indices, u.indices + u.k);
            std::sort(idx.begin(), idx.end());
            for (size_t j = 0; j < idx.size(); ++j)
                require(idx[j] < n && (j == 0 || idx[j] != idx[j - 1]), "update indices must be distinct and < n");
            written_fr.push_back({u.coeff_offset, u.coeff_offset + n});
            if (u.eval_offset != u.coeff_offset) written_fr.push_back({u.eval_offset, u.eval_offset + n});
            written_fr.push_back({u.value_slot, u.value_slot + 1});
            read_fr.push_back({u.linear_offset, u.linear_offset + n});
            written_pts.push_back({u.commitment_row * W, u.commitment_row * W + 2});
            written_pts.push_back({u.opening_cache_row * W, u.opening_cache_row * W + 2});
            read_pts.push_back({u.srs_row * W, (u.srs_row + n) * W});
            read_pts.push_back({u.opening_row * W, (u.opening_row + n) * W});
            entries += u.k;
            const uint64_t kk = u.k ? u.k : 1;  // an empty update runs a one-point MSM with scalar 0
            scalars += kk;
            msm_points += 2 * kk;
        }
        require(!sort_and_overlap(written_fr) && !sort_and_overlap(written_pts),
                "updates of one batch share session state");
        for (const Interval& x : read_fr) require(!hits(written_fr, x.lo, x.hi), "linear terms overlap written state");
        for (const Interval& x : read_pts) require(!hits(written_pts, x.lo, x.hi), "cached points overlap a table");
        require(msm_points < (1u << 28) && scalars < (1u << 28), "update batch too large");

        // ---- host packing: [table rows][local index][descs][scalars]
        const size_t m = count;
        const size_t idx_words = msm_points, loc_words = entries;
        const size_t desc_words = m * sizeof(UpdDesc) / 4, sc_words = scalars * 8;
        const size_t off_loc = idx_words;
        const size_t off_desc = (off_loc + loc_words + 3) & ~size_t(3);  // 16-byte alignment
        const size_t off_sc = (off_desc + desc_words + 3) & ~size_t(3);
        const size_t total_words = off_sc + sc_words;
        uint32_t* pin = o->pinned.as<uint32_t>(total_words);
        std::vector<MsmItem> items;
        items.reserve(2 * m);
        std::vector<NttBatchEntry> ntt;
        uint32_t* prow = pin;
        uint32_t* ploc = pin + off_loc;
        auto* pdesc = reinterpret_cast<UpdDesc*>(pin + off_desc);
        uint32_t* psc = pin + off_sc;
        uint32_t ipos = 0, epos = 0, spos = 0;
        uint32_t* frw = static_cast<uint32_t*>(fr);
        for (size_t r = 0; r < m; ++r) {
            const DvgpuUpdate& u = updates[r];
            const uint32_t k = u.k, kk = k ? k : 1;
            for (uint32_t j = 0; j < kk; ++j) {
                const uint32_t i = k ? u.indices[j] : 0;
                prow[ipos + j] = u.srs_row + i;
                prow[ipos + kk + j] = u.opening_row + i;
            }
            items.push_back({size_t(kk), ipos, spos, window_bits, uint32_t(W)});
            items.push_back({size_t(kk), ipos + kk, spos, window_bits, uint32_t(W)});
            if (k) {
                std::memcpy(ploc + epos, u.indices, size_t(k) * 4);
                std::memcpy(psc + size_t(spos) * 8, u.deltas, size_t(k) * 32);
            } else {
                std::memset(psc + size_t(spos) * 8, 0, 32);
            }
            pdesc[r] = {epos, k, spos, 0, u.coeff_offset, u.linear_offset, u.value_slot};
            ntt.push_back({frw + u.coeff_offset * 8, frw + u.eval_offset * 8, int(u.log_n), 1, false});
            ipos += 2 * kk;
            epos += k;
            spos += kk;
        }
        for (uint32_t t = 0; t < spos; ++t)
            require(is_canonical<FrParams>(*reinterpret_cast<const uint32_t(*)[8]>(psc + size_t(t) * 8)),
                    "update deltas must be canonical Fr");

        hipStream_t st = o->stream;
        uint32_t* dev = o->stage_a.as<uint32_t>(total_words);
        HIP_CHECK(hipMemcpyAsync(dev, pin, total_words * 4, hipMemcpyHostToDevice, st));
        // Plans (the MSM plan synchronises the stream before its upload, which
        // also completes the staged copy above; the NTT plan waits for its own
        // previous upload only).
        const WindowTable wt = update_window_table(o);
        MsmBatchPlan plan = msm_batch_plan(items, update_msm_config(o), &wt);
        plan.point_limit = uint32_t(point_count);
        for (size_t r = 0; r < m; ++r) {
            plan.descs[2 * r].result_slot = uint32_t(updates[r].commitment_row * W);
            plan.descs[2 * r + 1].result_slot = uint32_t(updates[r].opening_cache_row * W);
        }
        msm_batch_prepare(plan, PointLayout<FqOps>::XYZZ_WORDS, o->msm, st);
        ntt_domain_prepare(o->domain, rw, st);
        ntt_batch_prepare(o->domain, ntt, o->ntt_batch, st);
        uint32_t* win = o->msm.windows.as<uint32_t>(size_t(plan.total_windows) * PointLayout<FqOps>::XYZZ_WORDS);
        uint32_t* xyzz_out = o->msm.finals.as<uint32_t>(2 * m * 32);
        uint32_t* yout = o->stage_b.as<uint32_t>(m * 8);
        const uint32_t* d_rows = dev;
        const uint32_t* d_loc = dev + off_loc;
        const auto* d_desc = reinterpret_cast<const UpdDesc*>(dev + off_desc);
        const uint32_t* d_sc = dev + off_sc;
        auto* pts = static_cast<uint32_t*>(points);

        ScopedEvents<2> ev(device_ms != nullptr);
        if (device_ms) HIP_CHECK(hipEventRecord(ev[0], st));
        if (entries) {
            k_scatter_add<<<grid_for(entries, 256), 256, 0, st>>>(d_desc, uint32_t(m), uint32_t(entries), d_loc, d_sc,
                                                                  frw);
            HIP_CHECK_LAUNCH();
        }
        k_linear_dot<<<unsigned(m), 64, 0, st>>>(d_desc, d_loc, d_sc, frw, yout);
        HIP_CHECK_LAUNCH();
        msm_batch_enqueue<FqOps>(pts, d_rows, d_sc, plan, o->msm, st, win);
        k_add_to_cache<<<grid_for(2 * m, 64), 64, 0, st>>>(static_cast<const MsmDesc*>(o->msm.descs.ptr),
                                                           uint32_t(2 * m), win, pts,
                                                           static_cast<const uint32_t*>(o->msm.total.ptr) + 3,
                                                           xyzz_out);
        HIP_CHECK_LAUNCH();
        ntt_batch_enqueue(o->domain, o->ntt_batch, st);
        if (device_ms) HIP_CHECK(hipEventRecord(ev[1], st));
        uint8_t* host = o->msm.host_windows.as<uint8_t>(2 * m * 128 + m * 32 + 4);
        HIP_CHECK(hipMemcpyAsync(host, xyzz_out, 2 * m * 128, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(host + 2 * m * 128, yout, m * 32, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(host + 2 * m * 128 + m * 32, static_cast<uint32_t*>(o->msm.total.ptr) + 3, 4,
                                 hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (device_ms) *device_ms = ev.ms(0, 1);
        uint32_t flag = 0;
        std::memcpy(&flag, host + 2 * m * 128 + m * 32, 4);
        require(flag == 0, "MSM point index out of range");  // not reachable after the range checks
        // C and W of update r are XYZZ records 2r and 2r + 1.
        std::vector<uint8_t> affine(2 * m * 64);
        batch_to_affine(reinterpret_cast<const uint32_t*>(host), 2 * m, affine.data(), 64);
        for (size_t r = 0; r < m; ++r) {
            std::memcpy(results + 160 * r, affine.data() + 128 * r, 128);
            std::memcpy(results + 160 * r + 128, host + 2 * m * 128 + r * 32, 32);
        }
        return 0;
    });
}

int dvgpu_ext_update_reserve(void* owner, size_t max_updates, size_t max_entries, size_t max_elements,
                             int window_bits, const uint8_t* root, char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(max_updates && max_updates <= (1u << 20) && window_bits >= 4 && window_bits <= 20 && root &&
                    max_elements <= (size_t(1) << 31),
                "invalid reservation");
        uint32_t rw[8];
        std::memcpy(rw, root, 32);
        require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
        const size_t m = max_updates, e = std::max(max_entries, m);
        const uint32_t W = uint32_t((255 + window_bits - 1) / window_bits);
        // The largest plan a batch within these bounds can produce: 2m MSMs
        // over e points each side (the counts, not the split, size the buffers).
        std::vector<MsmItem> items(2 * m, MsmItem{1, 0, 0, window_bits, W});
        items[0].n = items[1].n = e - (m - 1);
        const WindowTable wt = update_window_table(o);
        MsmBatchPlan plan = msm_batch_plan(items, update_msm_config(o), &wt);
        hipStream_t st = o->stream;
        msm_batch_prepare(plan, PointLayout<FqOps>::XYZZ_WORDS, o->msm, st);
        o->msm.windows.as<uint32_t>(size_t(plan.total_windows) * PointLayout<FqOps>::XYZZ_WORDS);
        o->msm.finals.as<uint32_t>(2 * m * 32);
        o->msm.host_windows.as<uint8_t>(2 * m * 128 + m * 32 + 4);
        const size_t words = 2 * e + e + m * sizeof(UpdDesc) / 4 + e * 8 + 16;
        o->pinned.as<uint32_t>(words);
        o->stage_a.as<uint32_t>(words);
        o->stage_b.as<uint32_t>(m * 8);
        ntt_domain_prepare(o->domain, rw, st);
        o->ntt_batch.work.as<uint32_t>(std::max<size_t>(max_elements, 1) * 8);
        o->ntt_batch.descs.as<uint8_t>(m * 256);
        o->ntt_batch.host_descs.as<uint8_t>(m * 256);
        HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

}  // extern "C"
