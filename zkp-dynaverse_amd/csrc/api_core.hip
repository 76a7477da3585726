// C ABI: owner lifecycle, G1/G2 MSM, batched MSM and the Fr NTT.
// Entry-point semantics follow native/src/icicle_msm.cu (cited per function).
#include <algorithm>
#include <cstring>
#include <vector>

#include "owner.hpp"

namespace dvgpu {

int fail(int status, const char* message, char* error, size_t capacity) {
    if (error && capacity) {
        const size_t count = std::min(capacity - 1, std::strlen(message));
        std::memcpy(error, message, count);
        error[count] = '\0';
    }
    return status;
}

Owner::Owner() {
    msm.attach(&quota);
    ntt.attach(&quota);
    ntt_batch.attach(&quota);
    ntt_batch_inv.attach(&quota);
    domain.set_quota(&quota);
    stage_a.quota = stage_b.quota = stage_c.quota = &quota;
}

hipEvent_t* Owner::msm_phase_events() {
    if (!msm_phase[0]) {
        hipEvent_t made[MSM_PHASE_EVENTS] = {};
        for (int k = 0; k < MSM_PHASE_EVENTS; ++k) {
            const hipError_t st = hipEventCreate(&made[k]);
            if (st != hipSuccess) {
                for (int j = 0; j < k; ++j) (void)hipEventDestroy(made[j]);
                hip_check(st, "hipEventCreate");
            }
        }
        for (int k = 0; k < MSM_PHASE_EVENTS; ++k) msm_phase[k] = made[k];
    }
    return msm_phase;
}

void Owner::release_device() noexcept {
    if (resident) {
        retired_resident_stats = resident_stats(resident);
        retired_resident_diagnostics = resident_diagnostics(resident);
        resident_destroy(resident);
        resident = nullptr;
    }
    if (setup) {
        setup_destroy(setup);
        setup = nullptr;
    }
    msm.release();
    for (hipEvent_t& e : msm_phase) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    ntt.release();
    ntt_batch.release();
    ntt_batch_inv.release();
    poly_stage.release();
    domain.release();
    stage_a.release();
    stage_b.release();
    stage_c.release();
    pinned.release();
    for (BatchObj* b : batches) batch_destroy(b);
    batches.clear();
    for (KzgSrs* z : kzg) kzg_destroy(z);
    kzg.clear();
    for (PolyTree* t : poly_trees) poly_tree_destroy(t);
    poly_trees.clear();
    if (pairing) {
        pairing_cache_destroy(pairing);
        pairing = nullptr;
    }
    for (auto& kv : ext_allocations) {
        (void)hipFree(kv.first);
        quota.release(kv.second);
    }
    ext_allocations.clear();
}

void Owner::retire() noexcept {
    if (retired.exchange(true)) return;
    // Calls registered on other threads (KZG) finish before anything goes.
    gate.close_and_drain();
    if (stream) (void)hipStreamSynchronize(stream);
    release_device();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    (void)hipGetLastError();
}

Owner::~Owner() {
    if (!retired) {
        gate.close_and_drain();
        if (stream) (void)hipStreamSynchronize(stream);
        release_device();
        if (stream) (void)hipStreamDestroy(stream);
    }
}

namespace {

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

template <class F>
int msm_entry(Owner* o, const uint8_t* points, const uint8_t* scalars, size_t length, uint8_t* result,
              DvgpuMsmTimings* timings, DvgpuMsmTimeline* timeline, char* error, size_t capacity) {
    constexpr int PB = PointLayout<F>::CANON_BYTES;
    return guarded(o, error, capacity, [&]() -> int {
        require(result && length <= DVGPU_MAX_MSM_SIZE && (!length || (points && scalars)),
                "invalid MSM buffers or length");
        if (!length) {
            std::memset(result, 0, PB);
            return 0;
        }
        const uint64_t t0 = now_ns();
        uint8_t* dpts = o->stage_a.as<uint8_t>(length * PB);
        uint32_t* dsc = o->stage_b.as<uint32_t>(length * 8);
        uint32_t* mont = o->msm.points.as<uint32_t>(length * PointLayout<F>::AFFINE_WORDS);
        const uint64_t t1 = now_ns();
        ScopedEvents<MSM_PHASE_EVENTS + 2> ev(timeline != nullptr);
        if (timeline) HIP_CHECK(hipEventRecord(ev[MSM_PHASE_EVENTS], o->stream));
        HIP_CHECK(hipMemcpyAsync(dpts, points, length * PB, hipMemcpyHostToDevice, o->stream));
        HIP_CHECK(hipMemcpyAsync(dsc, scalars, length * 32, hipMemcpyHostToDevice, o->stream));
        points_to_mont<F>(dpts, mont, length, o->stream);
        if (timeline) {
            HIP_CHECK(hipEventRecord(ev[MSM_PHASE_EVENTS + 1], o->stream));
            o->msm.phase_events = ev.e;
        }
        Affine<F> r;
        MsmPlan plan;
        MsmConfig cfg = o->msm_config;
        if (!cfg.window_bits) cfg.window_bits = msm_window_for(length, &o->window_table);
        try {
            const size_t room = msm_plan(length, cfg).outputs;
            uint32_t* window_words = o->msm.windows.as<uint32_t>(room * PointLayout<F>::XYZZ_WORDS);
            plan = msm_enqueue<F>(mont, dsc, length, o->msm, o->stream, cfg, window_words, room);
            o->msm.phase_events = nullptr;
            uint32_t* host = o->msm.host_windows.as<uint32_t>(plan.outputs * PointLayout<F>::XYZZ_WORDS);
            HIP_CHECK(hipMemcpyAsync(host, window_words, plan.outputs * PointLayout<F>::XYZZ_WORDS * 4,
                                     hipMemcpyDeviceToHost, o->stream));
            HIP_CHECK(hipStreamSynchronize(o->stream));
        } catch (...) {
            o->msm.phase_events = nullptr;
            throw;
        }
        const uint64_t t2 = now_ns();
        r = msm_combine_windows<F>(o->msm.host_windows.as<uint32_t>(0), plan);
        uint8_t out[PB];
        affine_to_bytes<F>(r, out);
        std::memcpy(result, out, PB);
        const uint64_t t3 = now_ns();
        if (timings) {
            timings->conversion_ns = t1 - t0;
            timings->icicle_ns = t2 - t1;
            timings->output_ns = t3 - t2;
        }
        if (timeline) {
            // Phases: 0 upload+convert, 1 digits, 2 scan, 3 kernels (scatter,
            // tasks, accumulate, reduce, tree).  Kernel slots follow the
            // device phases in order.
            auto span = [&](int a, int b) { return ev.ns(a, b); };
            timeline->c = uint64_t(plan.c);
            timeline->bitsize = 254;
            timeline->windows = uint64_t(plan.windows);
            timeline->buckets = uint64_t(plan.buckets);
            timeline->event_ns[0] = span(MSM_PHASE_EVENTS, MSM_PHASE_EVENTS + 1);
            timeline->event_ns[1] = span(0, 1);
            timeline->event_ns[2] = span(1, 2);
            uint64_t kernels = 0;
            for (int k = 2; k < MSM_PHASE_EVENTS - 1; ++k) {
                const uint64_t v = span(k, k + 1);
                timeline->kernel_ns[k - 2] = v;
                timeline->kernel_calls[k - 2] = 1;
                kernels += v;
            }
            timeline->event_ns[3] = kernels;
            for (int p = 0; p < 4; ++p) timeline->event_calls[p] = 1;
            // Host clocks: the whole device phase is one wall span.
            timeline->wall_ns[0] = t2 - t1;
            timeline->calls[0] = 1;
            timeline->bytes[0] = uint64_t(length) * (PB + 32);
            timeline->instrumentation_inclusive_ns = 0;
            timeline->instrumentation_total_ns = 0;
        }
        return 0;
    });
}

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

extern "C" {

// icicle_msm.cu:188-221
int dvgpu_backend_create_with_large_bucket_factor(uint32_t factor, void** output, char* error,
                                                  size_t capacity) {
    if (!output) return fail(-1, "null backend output", error, capacity);
    *output = nullptr;
    if (factor > DVGPU_MAX_LARGE_BUCKET_FACTOR)
        return fail(-1, "large bucket factor must be within 0..=100", error, capacity);
    try {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            return fail(-2, "no visible HIP device", error, capacity);
        HIP_CHECK(hipSetDevice(0));
        hipDeviceProp_t props{};
        HIP_CHECK(hipGetDeviceProperties(&props, 0));
        auto owner = std::make_unique<Owner>();
        owner->large_bucket_factor = factor;
        owner->identity = std::string(props.name[0] ? props.name : "AMD Instinct") + "/" + props.gcnArchName;
        create_stream(owner->stream);
        *output = owner.release();
        return 0;
    } catch (const std::exception& e) {
        return fail(-2, e.what(), error, capacity);
    } catch (...) {
        return fail(-3, "unknown native exception", error, capacity);
    }
}

int dvgpu_backend_create(void** output, char* error, size_t capacity) {
    return dvgpu_backend_create_with_large_bucket_factor(0, output, error, capacity);
}

void dvgpu_backend_destroy(void* backend) { delete as_owner(backend); }

// icicle_msm.cu:231-238
int dvgpu_backend_device_identity(void* backend, char* output, size_t output_capacity, char* error,
                                  size_t capacity) {
    const Owner* o = as_owner(backend);
    if (!o || !output || output_capacity <= o->identity.size())
        return fail(-1, "invalid device identity output", error, capacity);
    std::memcpy(output, o->identity.c_str(), o->identity.size() + 1);
    return 0;
}

int dvgpu_backend_batch_stats(void* backend, DvgpuBatchStats* stats) {
    const Owner* o = as_owner(backend);
    if (!o || !stats) return -1;
    *stats = o->batch_stats;
    return 0;
}

// icicle_msm.cu:247-265
int dvgpu_backend_memory_info(void* backend, DvgpuMemoryInfo* info, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(info != nullptr, "null memory info output");
        size_t free_bytes = 0, total_bytes = 0;
        HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
        *info = {uint64_t(free_bytes), uint64_t(total_bytes)};
        return 0;
    });
}

// icicle_msm.cu:436-439
int dvgpu_bn254_msm(void* backend, const uint8_t* points, const uint8_t* scalars, size_t length,
                    uint8_t* result, char* error, size_t capacity) {
    return msm_entry<FqOps>(as_owner(backend), points, scalars, length, result, nullptr, nullptr,
                            error, capacity);
}

// icicle_msm.cu:882-888
int dvgpu_bn254_msm_profiled(void* backend, const uint8_t* points, const uint8_t* scalars,
                             size_t length, uint8_t* result, DvgpuMsmTimings* timings, char* error,
                             size_t capacity) {
    if (!timings) return fail(-1, "null diagnostic timings", error, capacity);
    *timings = {};
    return msm_entry<FqOps>(as_owner(backend), points, scalars, length, result, timings, nullptr,
                            error, capacity);
}

// icicle_msm.cu:890-897
int dvgpu_bn254_msm_timeline(void* backend, const uint8_t* points, const uint8_t* scalars,
                             size_t length, uint8_t* result, const char* label,
                             DvgpuMsmTimings* timings, DvgpuMsmTimeline* timeline, char* error,
                             size_t capacity) {
    if (!timings || !timeline || !label) return fail(-1, "null timeline diagnostics", error, capacity);
    *timings = {};
    *timeline = {};
    return msm_entry<FqOps>(as_owner(backend), points, scalars, length, result, timings, timeline,
                            error, capacity);
}

// icicle_msm.cu:441-472
int dvgpu_bn254_g2_msm(void* backend, const uint8_t* points, const uint8_t* scalars, size_t length,
                       uint8_t* result, char* error, size_t capacity) {
    return msm_entry<Fq2Ops>(as_owner(backend), points, scalars, length, result, nullptr, nullptr,
                             error, capacity);
}

// icicle_msm.cu:474-492
int dvgpu_reserve_batch(void* backend, size_t points, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(points <= DVGPU_MAX_MSM_SIZE, "invalid batch reserve capacity");
        if (points > o->batch_stats.input_capacity) {
            o->stage_a.as<uint8_t>(points * 64);
            o->stage_b.as<uint8_t>(points * 32);
            o->msm.points.as<uint32_t>(points * 16);
            ++o->batch_stats.buffer_grows;
            o->batch_stats.input_capacity = points;
        }
        return 0;
    });
}

// icicle_msm.cu:810-880: `count` MSMs of `length` points each.
int dvgpu_bn254_msm_batch(void* backend, const uint8_t* points, const uint8_t* scalars, size_t length,
                          size_t count, uint8_t* results, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(points && scalars && results && length && count && count <= DVGPU_MAX_MSM_BATCH_SIZE &&
                    length <= DVGPU_MAX_MSM_SIZE / count,
                "invalid batch buffers or bounded work size");
        const size_t total = length * count;
        ++o->batch_stats.icicle_calls;
        o->batch_stats.max_batch_size = std::max<uint64_t>(o->batch_stats.max_batch_size, count);
        if (total > o->batch_stats.input_capacity) {
            ++o->batch_stats.buffer_grows;
            o->batch_stats.input_capacity = total;
        }
        uint8_t* dpts = o->stage_a.as<uint8_t>(total * 64);
        uint32_t* dsc = o->stage_b.as<uint32_t>(total * 8);
        uint32_t* mont = o->msm.points.as<uint32_t>(total * 16);
        HIP_CHECK(hipMemcpyAsync(dpts, points, total * 64, hipMemcpyHostToDevice, o->stream));
        HIP_CHECK(hipMemcpyAsync(dsc, scalars, total * 32, hipMemcpyHostToDevice, o->stream));
        points_to_mont<FqOps>(dpts, mont, total, o->stream);
        // All sets in one batched kernel sequence; windows combined on the host pool.
        std::vector<MsmItem> items;
        for (size_t i = 0; i < count; ++i)
            items.push_back({length, uint32_t(i * length), uint32_t(i * length), 0});
        const MsmBatchPlan plan = msm_batch_plan(items, o->msm_config, &o->window_table);
        constexpr int XW = PointLayout<FqOps>::XYZZ_WORDS;
        msm_batch_prepare(plan, XW, o->msm, o->stream);
        uint32_t* win = o->msm.windows.as<uint32_t>(size_t(plan.total_windows) * XW);
        msm_batch_enqueue<FqOps>(mont, nullptr, dsc, plan, o->msm, o->stream, win);
        uint32_t* host = o->msm.host_windows.as<uint32_t>(size_t(plan.total_windows) * XW);
        HIP_CHECK(hipMemcpyAsync(host, win, size_t(plan.total_windows) * XW * 4, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        std::vector<Affine<FqOps>> res(count);
        msm_batch_combine<FqOps>(host, plan, res.data());
        uint8_t encoded[DVGPU_MAX_MSM_BATCH_SIZE * 64];
        for (size_t i = 0; i < count; ++i) affine_to_bytes<FqOps>(res[i], encoded + i * 64);
        std::memcpy(results, encoded, count * 64);
        return 0;
    });
}

// icicle_msm.cu:767-808
int dvgpu_bn254_ntt(void* backend, const uint8_t* values, size_t length, int inverse,
                    const uint8_t* root, uint8_t* result, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(values && result && root && length && length <= DVGPU_MAX_NTT_SIZE &&
                    !(length & (length - 1)) && (inverse == 0 || inverse == 1),
                "invalid NTT buffers, direction or length");
        uint32_t rw[8];
        std::memcpy(rw, root, 32);
        require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
        if (length == 1) {
            std::memmove(result, values, 32);
            return 0;
        }
        int log_n = 0;
        while ((size_t(1) << log_n) < length) ++log_n;
        ntt_domain_prepare(o->domain, rw, o->stream);
        uint32_t* din = o->stage_a.as<uint32_t>(length * 8);
        HIP_CHECK(hipMemcpyAsync(din, values, length * 32, hipMemcpyHostToDevice, o->stream));
        ntt_rows(o->domain, din, din, log_n, 1, inverse != 0, o->ntt, o->stream);
        uint8_t* staged = o->pinned.as<uint8_t>(length * 32);
        HIP_CHECK(hipMemcpyAsync(staged, din, length * 32, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        std::memcpy(result, staged, length * 32);
        return 0;
    });
}

// Pairing-tower probes (icicle_msm.cu:656-747) are diagnostics outside the
// proving path; they reject every call without touching the device.
int dvgpu_bn254_tower_inspect(void* owner, uint32_t, uint32_t, uint32_t, const uint8_t*, size_t,
                              const uint8_t*, size_t, uint8_t*, size_t, uint32_t*, size_t,
                              char* error, size_t capacity) {
    if (!owner) return fail(-4, "backend retired", error, capacity);
    return fail(-1, "BN254 tower diagnostics are not provided by this library", error, capacity);
}

int dvgpu_bn254_lines_inspect(void* owner, const uint8_t*, size_t, uint8_t*, size_t, uint32_t*,
                              size_t, uint32_t*, size_t, char* error, size_t capacity) {
    if (!owner) return fail(-4, "backend retired", error, capacity);
    return fail(-1, "BN254 line diagnostics are not provided by this library", error, capacity);
}

int dvgpu_bn254_miller_inspect(void* owner, const uint8_t*, size_t, const uint8_t*, size_t,
                               uint8_t*, size_t, uint32_t*, size_t, char* error, size_t capacity) {
    if (!owner) return fail(-4, "backend retired", error, capacity);
    return fail(-1, "BN254 Miller diagnostics are not provided by this library", error, capacity);
}

}  // extern "C"
