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
    NttDomain domain;
    NttWorkspace ntt_ws;
    const uint8_t* wire = nullptr;
    size_t wire_n = 0;
    uint32_t algorithm = 3;
    uint32_t* masks = nullptr;
    DeviceBuffer win1, win2, report_dev;
    HostBuffer host1, host2, report_host;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    void attach(Quota* q) {
        w1.attach(q);
        w2.attach(q);
        ntt_ws.attach(q);
        domain.fwd.quota = domain.inv.quota = q;
        win1.quota = win2.quota = report_dev.quota = q;
    }

    void enqueue(hipStream_t st) {
        if (!p1.descs.empty()) {
            uint32_t* win = static_cast<uint32_t*>(win1.ptr);
            msm_batch_enqueue<FqOps>(g1_points, g1_index, g1_scalars, p1, w1, st, win);
            HIP_CHECK(hipMemcpyAsync(host1.ptr, win, size_t(p1.total_windows) * PointLayout<FqOps>::XYZZ_WORDS * 4,
                                     hipMemcpyDeviceToHost, st));
        }
        if (!p2.descs.empty()) {
            uint32_t* win = static_cast<uint32_t*>(win2.ptr);
            msm_batch_enqueue<Fq2Ops>(g2_points, g2_index, g2_scalars, p2, w2, st, win);
            HIP_CHECK(hipMemcpyAsync(host2.ptr, win, size_t(p2.total_windows) * PointLayout<Fq2Ops>::XYZZ_WORDS * 4,
                                     hipMemcpyDeviceToHost, st));
        }
        for (const DvgpuNttRows& r : ntt) {
            int log_n = 0;
            while ((uint64_t(1) << log_n) < r.length) ++log_n;
            ntt_rows(domain, static_cast<const uint32_t*>(r.in), static_cast<uint32_t*>(r.out), log_n, r.rows,
                     r.inverse != 0, ntt_ws, st);
        }
        if (wire_n) {
            auto* rep = static_cast<DvgpuG2ValidationReport*>(report_dev.ptr);
            g2_validate_enqueue(wire, wire_n, algorithm, masks, rep, owner->g2_kernel.threads, st, nullptr, nullptr);
            HIP_CHECK(hipMemcpyAsync(report_host.ptr, rep, sizeof(DvgpuG2ValidationReport), hipMemcpyDeviceToHost, st));
        }
    }

    ~BatchObj() {
        if (owner && owner->stream) (void)hipStreamSynchronize(owner->stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        w1.release();
        w2.release();
        ntt_ws.release();
        domain.release();
    }
};

void batch_destroy(BatchObj* b) noexcept { delete b; }

namespace {

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

// Plan the non-empty items; map[i] is the caller's index of planned MSM i.
MsmBatchPlan plan_items(const DvgpuMsmItem* items, size_t count, const Owner* o, std::vector<size_t>& map) {
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
    return msm_batch_plan(v, o->msm_config, &o->window_table);
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

int dvgpu_ext_set_window_table(void* owner, const int32_t* table) {
    Owner* o = as_owner(owner);
    if (!o || o->retired) return -4;
    if (!table) return -1;
    for (int i = 0; i < 25; ++i)
        if (table[i] != 0 && (table[i] < 4 || table[i] > 20)) return -1;
    for (int i = 0; i < 25; ++i) o->window_table.c[i] = table[i];
    return 0;
}

int dvgpu_ext_msm_batch_device(void* owner, const void* points, const void* point_index, const void* scalars,
                               const DvgpuMsmItem* items, size_t count, int g2, uint8_t* results, float* phase_ms,
                               char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(items && results && count && (g2 == 0 || g2 == 1) && points && scalars, "invalid batch arguments");
        std::vector<size_t> map;
        const MsmBatchPlan plan = plan_items(items, count, o, map);
        const int xw = g2 ? PointLayout<Fq2Ops>::XYZZ_WORDS : PointLayout<FqOps>::XYZZ_WORDS;
        if (plan.descs.empty()) {
            std::memset(results, 0, count * (g2 ? 128 : 64));
            return 0;
        }
        msm_batch_prepare(plan, xw, o->msm, o->stream);
        uint32_t* win = o->msm.windows.as<uint32_t>(size_t(plan.total_windows) * xw);
        hipEvent_t ev[MSM_PHASE_EVENTS] = {};
        if (phase_ms) {
            for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
            o->msm.phase_events = ev;
        }
        try {
            if (g2)
                msm_batch_enqueue<Fq2Ops>(static_cast<const uint32_t*>(points), static_cast<const uint32_t*>(point_index),
                                          static_cast<const uint32_t*>(scalars), plan, o->msm, o->stream, win);
            else
                msm_batch_enqueue<FqOps>(static_cast<const uint32_t*>(points), static_cast<const uint32_t*>(point_index),
                                         static_cast<const uint32_t*>(scalars), plan, o->msm, o->stream, win);
        } catch (...) {
            o->msm.phase_events = nullptr;
            for (auto& e : ev)
                if (e) (void)hipEventDestroy(e);
            throw;
        }
        o->msm.phase_events = nullptr;
        uint32_t* host = o->msm.host_windows.as<uint32_t>(size_t(plan.total_windows) * xw);
        HIP_CHECK(hipMemcpyAsync(host, win, size_t(plan.total_windows) * xw * 4, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        if (phase_ms) {
            for (int k = 0; k + 1 < MSM_PHASE_EVENTS; ++k) HIP_CHECK(hipEventElapsedTime(&phase_ms[k], ev[k], ev[k + 1]));
            for (auto& e : ev) (void)hipEventDestroy(e);
        }
        if (g2) publish<Fq2Ops>(host, plan, map, count, results);
        else publish<FqOps>(host, plan, map, count, results);
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
            b->p1 = plan_items(spec->g1_items, spec->g1_count, o, b->map1);
            b->count1 = spec->g1_count;
            b->g1_points = static_cast<const uint32_t*>(spec->g1_points);
            b->g1_index = static_cast<const uint32_t*>(spec->g1_index);
            b->g1_scalars = static_cast<const uint32_t*>(spec->g1_scalars);
            if (!b->p1.descs.empty()) {
                msm_batch_prepare(b->p1, PointLayout<FqOps>::XYZZ_WORDS, b->w1, st);
                b->win1.as<uint32_t>(size_t(b->p1.total_windows) * PointLayout<FqOps>::XYZZ_WORDS);
                b->host1.as<uint32_t>(size_t(b->p1.total_windows) * PointLayout<FqOps>::XYZZ_WORDS);
            }
        }
        if (spec->g2_count) {
            require(spec->g2_items && spec->g2_points && spec->g2_scalars, "missing G2 batch buffers");
            b->p2 = plan_items(spec->g2_items, spec->g2_count, o, b->map2);
            b->count2 = spec->g2_count;
            b->g2_points = static_cast<const uint32_t*>(spec->g2_points);
            b->g2_index = static_cast<const uint32_t*>(spec->g2_index);
            b->g2_scalars = static_cast<const uint32_t*>(spec->g2_scalars);
            if (!b->p2.descs.empty()) {
                msm_batch_prepare(b->p2, PointLayout<Fq2Ops>::XYZZ_WORDS, b->w2, st);
                b->win2.as<uint32_t>(size_t(b->p2.total_windows) * PointLayout<Fq2Ops>::XYZZ_WORDS);
                b->host2.as<uint32_t>(size_t(b->p2.total_windows) * PointLayout<Fq2Ops>::XYZZ_WORDS);
            }
        }
        if (spec->ntt_count) {
            require(spec->ntt && spec->root, "missing NTT batch buffers or root");
            uint32_t rw[8];
            std::memcpy(rw, spec->root, 32);
            require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
            for (size_t i = 0; i < spec->ntt_count; ++i) {
                const DvgpuNttRows& r = spec->ntt[i];
                require(r.in && r.out && r.length && r.length <= DVGPU_MAX_NTT_SIZE && !(r.length & (r.length - 1)) &&
                            r.rows && r.length * r.rows <= DVGPU_MAX_TRANSFORM_MATRIX_SIZE * 8ull,
                        "invalid NTT row batch");
                b->ntt.push_back(r);
            }
            ntt_domain_prepare(b->domain, rw, st);
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
        HIP_CHECK(hipEventCreate(&b->e0));
        HIP_CHECK(hipEventCreate(&b->e1));
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
        if (g1_results && b->count1)
            publish<FqOps>(static_cast<const uint32_t*>(b->host1.ptr), b->p1, b->map1, b->count1, g1_results);
        if (g2_results && b->count2)
            publish<Fq2Ops>(static_cast<const uint32_t*>(b->host2.ptr), b->p2, b->map2, b->count2, g2_results);
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
