// Batching form: many MSMs of uneven sizes (G1 and G2), NTT row batches and
// G2 validation run as one fixed launch sequence on the owner's stream,
// optionally captured once into a hipGraph and replayed with one launch.
//
// A batch owns its plans, workspaces, NTT domain and result buffers, so a
// replay reads nothing that later foreground calls might reallocate.  The
// sequence contains only kernels and copies (fill kernels instead of memsets;
// every buffer is sized and the descriptors uploaded before capture).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/dynaverse_hip_ext.h"
#include "owner.hpp"
#include "setup.hpp"

namespace dvgpu {

struct BatchObj {
    Owner* owner = nullptr;
    // MSMs: index into the caller's item list for every planned MSM.
    MsmBatchPlan p1, p2;
    std::vector<size_t> map1, map2;
    size_t count1 = 0, count2 = 0;
    MsmWorkspace w1, w2;
    const uint32_t *g1_points = nullptr, *g1_index = nullptr, *g1_scalars = nullptr;
    const uint32_t *g2_points = nullptr, *g2_index = nullptr, *g2_scalars = nullptr;
    std::vector<DvgpuNttRows> ntt;
    // Consecutive NTT entries that do not overlap run as one fused group.
    std::vector<std::unique_ptr<NttBatchPlan>> ntt_groups;
    NttDomain domain;
    const uint8_t* wire = nullptr;
    size_t wire_n = 0;
    uint32_t algorithm = 3;
    uint32_t* masks = nullptr;
    DeviceBuffer win1, win2, fin1, fin2, report_dev;
    HostBuffer host1, host2, report_host;
    bool dev_fin1 = false, dev_fin2 = false;  // window sums finished on the device
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    void attach(Quota* q) {
        w1.attach(q);
        w2.attach(q);
        domain.set_quota(q);
        win1.quota = win2.quota = fin1.quota = fin2.quota = report_dev.quota = q;
    }

    template <class F>
    void enqueue_msms(const uint32_t* pts, const uint32_t* idx, const uint32_t* sc, const MsmBatchPlan& p,
                      MsmWorkspace& w, DeviceBuffer& win, DeviceBuffer& fin, HostBuffer& host, bool dev_fin,
                      size_t count, hipStream_t st);

    // NTT rows and G2 validation share no buffer with the MSMs: they run on a
    // second stream forked from and joined back into the owner's stream (in a
    // captured graph: a parallel branch), so the latency-bound subgroup
    // checks overlap the MSM kernels.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;

    void enqueue(hipStream_t st) {
        const bool branch = wire_n || !ntt_groups.empty();
        if (branch) {
            HIP_CHECK(hipEventRecord(fork, st));
            HIP_CHECK(hipStreamWaitEvent(side, fork, 0));
            for (const auto& g : ntt_groups) ntt_batch_enqueue(domain, *g, side);
            if (wire_n) {
                auto* rep = static_cast<DvgpuG2ValidationReport*>(report_dev.ptr);
                G2CheckLaunch launch;  // the batch holds no repack scratch: records stay in place
                launch.config = owner->g2_kernel;
                g2_validate_enqueue(wire, wire_n, algorithm, masks, rep, launch, side);
                HIP_CHECK(hipMemcpyAsync(report_host.ptr, rep, sizeof(DvgpuG2ValidationReport),
                                         hipMemcpyDeviceToHost, side));
            }
            HIP_CHECK(hipEventRecord(join, side));
        }
        if (!p1.descs.empty())
            enqueue_msms<FqOps>(g1_points, g1_index, g1_scalars, p1, w1, win1, fin1, host1, dev_fin1, count1, st);
        if (!p2.descs.empty())
            enqueue_msms<Fq2Ops>(g2_points, g2_index, g2_scalars, p2, w2, win2, fin2, host2, dev_fin2, count2, st);
        if (branch) HIP_CHECK(hipStreamWaitEvent(st, join, 0));
    }

    ~BatchObj() {
        if (owner && owner->stream) (void)hipStreamSynchronize(owner->stream);
        if (side) (void)hipStreamSynchronize(side);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
        w1.release();
        w2.release();
        for (auto& g : ntt_groups) g->release();
        domain.release();
    }
};

void batch_destroy(BatchObj* b) noexcept { delete b; }

template <class F>
void BatchObj::enqueue_msms(const uint32_t* pts, const uint32_t* idx, const uint32_t* sc, const MsmBatchPlan& p,
                            MsmWorkspace& w, DeviceBuffer& win, DeviceBuffer& fin, HostBuffer& host, bool dev_fin,
                            size_t count, hipStream_t st) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS, PB = PointLayout<F>::CANON_BYTES;
    uint32_t* wv = static_cast<uint32_t*>(win.ptr);
    msm_batch_enqueue<F>(pts, idx, sc, p, w, st, wv);
    if (dev_fin) {
        uint8_t* out = static_cast<uint8_t*>(fin.ptr);
        device_fill_words(reinterpret_cast<uint32_t*>(out), count * PB / 4, 0u, st);
        msm_batch_finalize_enqueue<F>(wv, p, w, out, nullptr, st);
        HIP_CHECK(hipMemcpyAsync(host.ptr, out, count * PB, hipMemcpyDeviceToHost, st));
    } else {
        HIP_CHECK(hipMemcpyAsync(host.ptr, wv, size_t(p.total_windows) * XW * 4, hipMemcpyDeviceToHost, st));
    }
}

// Plan the non-empty items; map[i] is the caller's index of planned MSM i.
// G2 MSMs: every group operation costs about three G1 ones, so the serial
// bucket-reduction chains are cut short (two buckets per reduction thread,
// whatever the G1-tuned per-size table says).
static MsmConfig g2_config(const Owner* o, WindowTable& table) {
    MsmConfig cfg = o->msm_config;
    cfg.chunk = std::min(cfg.chunk, 2);
    table = o->window_table;
    for (int& ch : table.chunk) ch = 0;
    return cfg;
}

MsmBatchPlan plan_items(const DvgpuMsmItem* items, size_t count, const Owner* o, std::vector<size_t>& map,
                        bool g2) {
    std::vector<MsmItem> v;
    for (size_t i = 0; i < count; ++i) {
        require(items[i].n <= DVGPU_MAX_MSM_SIZE * 16ull, "batched MSM length out of range");
        require(items[i].window_bits == 0 || (items[i].window_bits >= 4 && items[i].window_bits <= 20),
                "batched MSM window bits must be 0 or within 4..20");
        if (!items[i].n) continue;
        map.push_back(i);
        v.push_back({size_t(items[i].n), items[i].point_base, items[i].scalar_base,
                     items[i].window_bits ? items[i].window_bits
                                          : msm_window_for(size_t(items[i].n), &o->window_table)});
    }
    if (v.empty()) return MsmBatchPlan{};
    WindowTable g2_table;
    const MsmConfig cfg = g2 ? g2_config(o, g2_table) : o->msm_config;
    MsmBatchPlan plan = msm_batch_plan(v, cfg, g2 ? &g2_table : &o->window_table);
    for (size_t k = 0; k < plan.descs.size(); ++k) plan.descs[k].out_index = uint32_t(map[k]);
    return plan;
}

namespace {

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

// Batches of at least this many MSMs finish their window sums on the device.
size_t finalize_threshold(const Owner* o) { return o->msm_finalize_min > 0 ? size_t(o->msm_finalize_min) : 4096; }

// Enqueue the device epilogue: zero every result (MSMs skipped as empty stay
// the identity), then one lane per planned MSM.
template <class F>
void finalize_enqueue(const uint32_t* win, const MsmBatchPlan& plan, const MsmWorkspace& ws, size_t count,
                      uint8_t* out, uint32_t* slot_points, hipStream_t st) {
    device_fill_words(reinterpret_cast<uint32_t*>(out), count * PointLayout<F>::CANON_BYTES / 4, 0u, st);
    msm_batch_finalize_enqueue<F>(win, plan, ws, out, slot_points, st);
}

NttBatchEntry ntt_entry(const DvgpuNttRows& r) {
    require(r.in && r.out && r.length && r.length <= DVGPU_MAX_NTT_SIZE && !(r.length & (r.length - 1)) && r.rows &&
                r.length * r.rows <= DVGPU_MAX_TRANSFORM_MATRIX_SIZE * 8ull,
            "invalid NTT row batch");
    int log_n = 0;
    while ((uint64_t(1) << log_n) < r.length) ++log_n;
    return {static_cast<const uint32_t*>(r.in), static_cast<uint32_t*>(r.out), log_n, size_t(r.rows), r.inverse != 0};
}

bool ntt_overlaps(const NttBatchEntry& x, const NttBatchEntry& y) {
    const uintptr_t xb = (x.rows << x.log_n) * 32, yb = (y.rows << y.log_n) * 32;
    auto hit = [](uintptr_t a, uintptr_t an, uintptr_t b, uintptr_t bn) { return a < b + bn && b < a + an; };
    const uintptr_t xi = reinterpret_cast<uintptr_t>(x.in), xo = reinterpret_cast<uintptr_t>(x.out);
    const uintptr_t yi = reinterpret_cast<uintptr_t>(y.in), yo = reinterpret_cast<uintptr_t>(y.out);
    return hit(xo, xb, yi, yb) || hit(xo, xb, yo, yb) || hit(yo, yb, xi, xb);
}

template <class F>
void publish(const uint32_t* words, const MsmBatchPlan& plan, const std::vector<size_t>& map, size_t count,
             uint8_t* results) {
    constexpr int PB = PointLayout<F>::CANON_BYTES;
    std::vector<uint8_t> out(count * PB, 0);
    if (!plan.descs.empty()) {
        std::vector<Affine<F>> pts(plan.descs.size());
        msm_batch_combine<F>(words, plan, pts.data());
        for (size_t k = 0; k < pts.size(); ++k) affine_to_bytes<F>(pts[k], out.data() + map[k] * PB);
    }
    std::memcpy(results, out.data(), count * PB);
}

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

extern "C" {

int dvgpu_ext_set_tuning(void* owner, const DvgpuTuning* t) {
    Owner* o = as_owner(owner);
    if (!o || o->retired) return -4;
    if (!t) return -1;
    for (int32_t c : t->msm_window_bits)
        if (c != 0 && (c < 4 || c > 20)) return -1;
    for (int32_t s : t->ntt_pass_stages)
        if (s < 0 || s > 11) return -1;
    if (t->msm_task_points < 0 || t->msm_task_points > 1024) return -1;
    if (t->msm_device_finalize_min < 0) return -1;
    if (t->msm_chunk != 0 && t->msm_chunk != 1 && t->msm_chunk != 2 && t->msm_chunk != 4 && t->msm_chunk != 8)
        return -1;
    for (int32_t ch : t->msm_chunk_by_log2)
        if (ch != 0 && ch != 1 && ch != 2 && ch != 4 && ch != 8) return -1;
    for (int i = 0; i < 25; ++i) o->window_table.c[i] = t->msm_window_bits[i];
    for (int i = 0; i < 25; ++i) o->window_table.chunk[i] = t->msm_chunk_by_log2[i];
    for (int i = 0; i <= NTT_MAX_LOG; ++i) o->domain.pass_stages[i] = t->ntt_pass_stages[i];
    o->msm_config.task_points = t->msm_task_points ? t->msm_task_points : MsmConfig{}.task_points;
    o->msm_config.chunk = t->msm_chunk ? t->msm_chunk : MsmConfig{}.chunk;
    o->msm_finalize_min = t->msm_device_finalize_min;
    return 0;
}

int dvgpu_ext_msm_window(void* owner, uint64_t n) {
    Owner* o = as_owner(owner);
    if (!o || o->retired) return -4;
    if (!n) return -1;
    return o->msm_config.window_bits ? o->msm_config.window_bits : msm_window_for(size_t(n), &o->window_table);
}

int dvgpu_ext_get_tuning(void* owner, DvgpuTuning* t) {
    Owner* o = as_owner(owner);
    if (!o || o->retired) return -4;
    if (!t) return -1;
    for (int i = 0; i < 25; ++i) t->msm_window_bits[i] = o->window_table.c[i];
    for (int i = 0; i < 25; ++i) t->msm_chunk_by_log2[i] = o->window_table.chunk[i];
    for (int i = 0; i <= NTT_MAX_LOG; ++i) t->ntt_pass_stages[i] = o->domain.pass_stages[i];
    t->msm_task_points = o->msm_config.task_points;
    t->msm_chunk = o->msm_config.chunk;
    t->msm_device_finalize_min = o->msm_finalize_min;
    return 0;
}

int dvgpu_ext_msm_batch_device(void* owner, const void* points, const void* point_index, const void* scalars,
                               const DvgpuMsmItem* items, size_t count, int g2, uint8_t* results, float* phase_ms,
                               char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(items && results && count && (g2 == 0 || g2 == 1) && points && scalars, "invalid batch arguments");
        std::vector<size_t> map;
        const MsmBatchPlan plan = plan_items(items, count, o, map, g2 != 0);
        const int xw = g2 ? PointLayout<Fq2Ops>::XYZZ_WORDS : PointLayout<FqOps>::XYZZ_WORDS;
        if (plan.descs.empty()) {
            std::memset(results, 0, count * (g2 ? 128 : 64));
            return 0;
        }
        msm_batch_prepare(plan, xw, o->msm, o->stream);
        uint32_t* win = o->msm.windows.as<uint32_t>(size_t(plan.total_windows) * xw);
        ScopedEvents<MSM_PHASE_EVENTS> ev(phase_ms != nullptr);
        if (phase_ms) o->msm.phase_events = ev.e;
        try {
            if (g2)
                msm_batch_enqueue<Fq2Ops>(static_cast<const uint32_t*>(points), static_cast<const uint32_t*>(point_index),
                                          static_cast<const uint32_t*>(scalars), plan, o->msm, o->stream, win);
            else
                msm_batch_enqueue<FqOps>(static_cast<const uint32_t*>(points), static_cast<const uint32_t*>(point_index),
                                         static_cast<const uint32_t*>(scalars), plan, o->msm, o->stream, win);
        } catch (...) {
            o->msm.phase_events = nullptr;
            throw;
        }
        o->msm.phase_events = nullptr;
        const bool on_device = plan.descs.size() >= finalize_threshold(o);
        const size_t pb = g2 ? 128 : 64;
        uint32_t* host = nullptr;
        if (on_device) {
            uint8_t* fin = o->msm.finals.as<uint8_t>(count * pb);
            if (g2) finalize_enqueue<Fq2Ops>(win, plan, o->msm, count, fin, nullptr, o->stream);
            else finalize_enqueue<FqOps>(win, plan, o->msm, count, fin, nullptr, o->stream);
            host = o->msm.host_windows.as<uint32_t>(count * pb / 4);
            HIP_CHECK(hipMemcpyAsync(host, fin, count * pb, hipMemcpyDeviceToHost, o->stream));
        } else {
            host = o->msm.host_windows.as<uint32_t>(size_t(plan.total_windows) * xw);
            HIP_CHECK(hipMemcpyAsync(host, win, size_t(plan.total_windows) * xw * 4, hipMemcpyDeviceToHost,
                                     o->stream));
        }
        HIP_CHECK(hipStreamSynchronize(o->stream));
        if (phase_ms)
            for (int k = 0; k + 1 < MSM_PHASE_EVENTS; ++k) phase_ms[k] = ev.ms(k, k + 1);
        if (on_device) std::memcpy(results, host, count * pb);
        else if (g2) publish<Fq2Ops>(host, plan, map, count, results);
        else publish<FqOps>(host, plan, map, count, results);
        return 0;
    });
}

int dvgpu_ext_msm_batch_finalize(void* owner, const void* points, size_t point_count, const void* point_index,
                                 size_t index_count, const void* scalars, size_t scalar_count,
                                 const DvgpuMsmItem* items, size_t count, int g2, void* results_device,
                                 uint8_t* results_host, const uint32_t* result_slots, void* slot_points,
                                 size_t slot_count, char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(items && count && (g2 == 0 || g2 == 1) && points && point_count && scalars &&
                    (results_device || results_host),
                "invalid batch arguments");
        require(point_count <= 0xffffffffull, "too many points");
        require(!result_slots == !slot_points, "result slots and slot points go together");
        for (size_t i = 0; i < count; ++i) {
            const DvgpuMsmItem& it = items[i];
            require(uint64_t(it.scalar_base) + it.n <= scalar_count, "MSM scalars out of range");
            if (point_index) require(uint64_t(it.point_base) + it.n <= index_count, "MSM point indices out of range");
            else require(uint64_t(it.point_base) + it.n <= point_count, "MSM points out of range");
            if (result_slots && result_slots[i] != MSM_NO_SLOT) {
                require(result_slots[i] < slot_count, "result slot out of range");
                require(it.n, "an empty MSM cannot store a result slot");
            }
        }
        std::vector<size_t> map;
        MsmBatchPlan plan = plan_items(items, count, o, map, g2 != 0);
        plan.point_limit = uint32_t(point_count);
        const size_t pb = g2 ? 128 : 64;
        uint8_t* out = results_device ? static_cast<uint8_t*>(results_device) : o->msm.finals.as<uint8_t>(count * pb);
        uint32_t* host = o->msm.host_windows.as<uint32_t>(count * pb / 4 + 1);
        if (plan.descs.empty()) {
            device_fill_words(reinterpret_cast<uint32_t*>(out), count * pb / 4, 0u, o->stream);
            host[count * pb / 4] = 0;
        } else {
            if (result_slots)
                for (auto& d : plan.descs) d.result_slot = result_slots[d.out_index];
            const int xw = g2 ? PointLayout<Fq2Ops>::XYZZ_WORDS : PointLayout<FqOps>::XYZZ_WORDS;
            msm_batch_prepare(plan, xw, o->msm, o->stream);
            uint32_t* win = o->msm.windows.as<uint32_t>(size_t(plan.total_windows) * xw);
            const auto* pts = static_cast<const uint32_t*>(points);
            const auto* idx = static_cast<const uint32_t*>(point_index);
            const auto* sc = static_cast<const uint32_t*>(scalars);
            auto* slots = static_cast<uint32_t*>(slot_points);
            if (g2) {
                msm_batch_enqueue<Fq2Ops>(pts, idx, sc, plan, o->msm, o->stream, win);
                finalize_enqueue<Fq2Ops>(win, plan, o->msm, count, out, slots, o->stream);
            } else {
                msm_batch_enqueue<FqOps>(pts, idx, sc, plan, o->msm, o->stream, win);
                finalize_enqueue<FqOps>(win, plan, o->msm, count, out, slots, o->stream);
            }
            HIP_CHECK(hipMemcpyAsync(host + count * pb / 4, static_cast<uint32_t*>(o->msm.total.ptr) + 3, 4,
                                     hipMemcpyDeviceToHost, o->stream));
        }
        if (results_host) HIP_CHECK(hipMemcpyAsync(host, out, count * pb, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        // The finalize kernel stored no slot if the error word was set.
        require(host[count * pb / 4] == 0, "MSM point index out of range");
        if (results_host) std::memcpy(results_host, host, count * pb);
        return 0;
    });
}

int dvgpu_ext_ntt_batch_device(void* owner, const DvgpuNttRows* rows, size_t count, const uint8_t* root,
                               float* kernel_ms, char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(rows && count && root, "invalid NTT batch arguments");
        uint32_t rw[8];
        std::memcpy(rw, root, 32);
        require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
        std::vector<NttBatchEntry> entries;
        for (size_t i = 0; i < count; ++i) entries.push_back(ntt_entry(rows[i]));
        ntt_domain_prepare(o->domain, rw, o->stream);
        ntt_batch_prepare(o->domain, entries, o->ntt_batch, o->stream);
        ScopedEvents<2> ev(kernel_ms != nullptr);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[0], o->stream));
        ntt_batch_enqueue(o->domain, o->ntt_batch, o->stream);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[1], o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        if (kernel_ms) *kernel_ms = ev.ms(0, 1);
        return 0;
    });
}

int dvgpu_ext_batch_create(void* owner, const DvgpuBatchSpec* spec, int use_graph, void** out, char* error,
                           size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(spec && out, "invalid batch specification");
        auto b = std::make_unique<BatchObj>();
        b->owner = o;
        b->attach(&o->quota);
        b->use_graph = use_graph != 0;
        hipStream_t st = o->stream;
        if (spec->g1_count) {
            require(spec->g1_items && spec->g1_points && spec->g1_scalars, "missing G1 batch buffers");
            b->p1 = plan_items(spec->g1_items, spec->g1_count, o, b->map1, false);
            b->count1 = spec->g1_count;
            b->g1_points = static_cast<const uint32_t*>(spec->g1_points);
            b->g1_index = static_cast<const uint32_t*>(spec->g1_index);
            b->g1_scalars = static_cast<const uint32_t*>(spec->g1_scalars);
            if (!b->p1.descs.empty()) {
                msm_batch_prepare(b->p1, PointLayout<FqOps>::XYZZ_WORDS, b->w1, st);
                b->win1.as<uint32_t>(size_t(b->p1.total_windows) * PointLayout<FqOps>::XYZZ_WORDS);
                b->dev_fin1 = b->p1.descs.size() >= finalize_threshold(o);
                if (b->dev_fin1) {
                    b->fin1.as<uint8_t>(b->count1 * 64);
                    b->host1.as<uint8_t>(b->count1 * 64);
                } else {
                    b->host1.as<uint32_t>(size_t(b->p1.total_windows) * PointLayout<FqOps>::XYZZ_WORDS);
                }
            }
        }
        if (spec->g2_count) {
            require(spec->g2_items && spec->g2_points && spec->g2_scalars, "missing G2 batch buffers");
            b->p2 = plan_items(spec->g2_items, spec->g2_count, o, b->map2, true);
            b->count2 = spec->g2_count;
            b->g2_points = static_cast<const uint32_t*>(spec->g2_points);
            b->g2_index = static_cast<const uint32_t*>(spec->g2_index);
            b->g2_scalars = static_cast<const uint32_t*>(spec->g2_scalars);
            if (!b->p2.descs.empty()) {
                msm_batch_prepare(b->p2, PointLayout<Fq2Ops>::XYZZ_WORDS, b->w2, st);
                b->win2.as<uint32_t>(size_t(b->p2.total_windows) * PointLayout<Fq2Ops>::XYZZ_WORDS);
                b->dev_fin2 = b->p2.descs.size() >= finalize_threshold(o);
                if (b->dev_fin2) {
                    b->fin2.as<uint8_t>(b->count2 * 128);
                    b->host2.as<uint8_t>(b->count2 * 128);
                } else {
                    b->host2.as<uint32_t>(size_t(b->p2.total_windows) * PointLayout<Fq2Ops>::XYZZ_WORDS);
                }
            }
        }
        if (spec->ntt_count) {
            require(spec->ntt && spec->root, "missing NTT batch buffers or root");
            uint32_t rw[8];
            std::memcpy(rw, spec->root, 32);
            require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
            std::memcpy(b->domain.pass_stages, o->domain.pass_stages, sizeof(b->domain.pass_stages));
            b->domain.set_quota(&o->quota);
            ntt_domain_prepare(b->domain, rw, st);
            // Entries run in order; consecutive entries with no overlap fuse.
            std::vector<std::vector<NttBatchEntry>> groups;
            for (size_t i = 0; i < spec->ntt_count; ++i) {
                b->ntt.push_back(spec->ntt[i]);
                const NttBatchEntry e = ntt_entry(spec->ntt[i]);
                bool fits = !groups.empty();
                if (fits)
                    for (const NttBatchEntry& g : groups.back())
                        if (ntt_overlaps(e, g)) fits = false;
                if (!fits) groups.emplace_back();
                groups.back().push_back(e);
            }
            for (const auto& g : groups) {
                b->ntt_groups.push_back(std::make_unique<NttBatchPlan>());
                b->ntt_groups.back()->attach(&o->quota);
                ntt_batch_prepare(b->domain, g, *b->ntt_groups.back(), st);
            }
        }
        if (spec->g2_wire_n) {
            require(spec->g2_wire && spec->g2_masks && spec->g2_wire_n <= (1u << 20) && spec->g2_algorithm >= 1 &&
                        spec->g2_algorithm <= 3,
                    "invalid G2 validation batch");
            b->wire = static_cast<const uint8_t*>(spec->g2_wire);
            b->wire_n = spec->g2_wire_n;
            b->algorithm = spec->g2_algorithm;
            b->masks = static_cast<uint32_t*>(spec->g2_masks);
            b->report_dev.as<DvgpuG2ValidationReport>(1);
            b->report_host.as<DvgpuG2ValidationReport>(1);
        }
        create_event(b->e0);
        create_event(b->e1);
        create_event(b->fork, hipEventDisableTiming);
        create_event(b->join, hipEventDisableTiming);
        create_stream(b->side);
        // Warm-up run: sizes every remaining workspace before any capture.
        b->enqueue(st);
        HIP_CHECK(hipStreamSynchronize(st));
        if (b->use_graph) {
            HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                b->enqueue(st);
            } catch (...) {
                hipGraph_t discard = nullptr;
                (void)hipStreamEndCapture(st, &discard);
                if (discard) (void)hipGraphDestroy(discard);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(st, &b->graph));
            HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        o->batches.push_back(b.get());
        *out = b.release();
        return 0;
    });
}

int dvgpu_ext_batch_run(void* owner, void* batch, uint8_t* g1_results, uint8_t* g2_results,
                        DvgpuG2ValidationReport* report, float* device_ms, char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        auto* b = static_cast<BatchObj*>(batch);
        require(b && std::find(o->batches.begin(), o->batches.end(), b) != o->batches.end(),
                "unknown batch for this owner");
        hipStream_t st = o->stream;
        HIP_CHECK(hipEventRecord(b->e0, st));
        if (b->use_graph) HIP_CHECK(hipGraphLaunch(b->exec, st));
        else b->enqueue(st);
        HIP_CHECK(hipEventRecord(b->e1, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (device_ms) HIP_CHECK(hipEventElapsedTime(device_ms, b->e0, b->e1));
        if (g1_results && b->count1) {
            if (b->dev_fin1) std::memcpy(g1_results, b->host1.ptr, b->count1 * 64);
            else publish<FqOps>(static_cast<const uint32_t*>(b->host1.ptr), b->p1, b->map1, b->count1, g1_results);
        }
        if (g2_results && b->count2) {
            if (b->dev_fin2) std::memcpy(g2_results, b->host2.ptr, b->count2 * 128);
            else publish<Fq2Ops>(static_cast<const uint32_t*>(b->host2.ptr), b->p2, b->map2, b->count2, g2_results);
        }
        if (report && b->wire_n) std::memcpy(report, b->report_host.ptr, sizeof(DvgpuG2ValidationReport));
        return 0;
    });
}

int dvgpu_ext_batch_destroy(void* owner, void* batch) {
    Owner* o = as_owner(owner);
    if (!o) return -1;
    auto it = std::find(o->batches.begin(), o->batches.end(), static_cast<BatchObj*>(batch));
    if (it == o->batches.end()) return -1;
    o->batches.erase(it);
    batch_destroy(static_cast<BatchObj*>(batch));
    return 0;
}

}  // extern "C"
