// The opaque owner behind every `void* backend` of the C ABI.
//
// The reference binds one isolated CUDA context per owner
// (native/src/icicle_msm.cu:92-144).  HIP has no user-visible contexts, so an
// owner here is: device 0, one non-blocking stream, its own grow-only device
// scratch and domains, and a retired flag.  Any device error retires the
// owner (all its device memory is released and later calls return -4);
// invalid arguments never do.
#pragma once
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/dynaverse_gpu.h"
#include "../../include/dynaverse_hip_ext.h"
#include "call_gate.hpp"
#include "common.hpp"
#include "msm.hpp"
#include "ntt.hpp"

namespace dvgpu {

struct ResidentState;  // resident.hip
struct BatchObj;       // batch.hip
struct SetupState;     // setup.hip
struct KzgSrs;         // kzg.hip
struct PolyTree;       // poly.hip
struct PairingCache;   // pairing.hip

struct Owner {
    hipStream_t stream = nullptr;
    std::atomic<bool> retired{false};
    // KZG calls (several threads per SRS allowed) register here; retire()
    // closes the gate and waits for them before releasing anything.
    CallGate gate;
    uint32_t large_bucket_factor = 0;
    std::string identity;
    Quota quota;                // limit 0 until the resident engine is configured
    MsmWorkspace msm;
    NttWorkspace ntt;
    NttDomain domain;
    NttBatchPlan ntt_batch;                  // fused NTT batches (dynaverse_hip_ext.h)
    NttBatchPlan ntt_batch_inv;              // inverse half of a polynomial product round (dynaverse_poly.h)
    HostBuffer poly_stage;                   // pinned operands of a polynomial product round
    DeviceBuffer stage_a, stage_b, stage_c;  // per-call device inputs/outputs
    HostBuffer pinned;                       // per-call host staging
    DvgpuBatchStats batch_stats{};
    MsmConfig msm_config;
    // Fixed-base setup and G2 validation state.
    SetupState* setup = nullptr;
    DvgpuG2ValidationStats g2_stats{};
    DvgpuG2ValidationTraffic g2_traffic{};
    DvgpuG2KernelConfig g2_kernel{128, 0, 0, 0};
    DvgpuG2ExperimentStats g2_experiment{};
    int64_t g2_corrupt_next = -1;
    // Resident operator engine.
    ResidentState* resident = nullptr;
    DvgpuResidentStats retired_resident_stats{};
    DvgpuResidentDiagnostics retired_resident_diagnostics{};
    // Caller-visible device allocations (dynaverse_hip_ext.h).
    std::unordered_map<void*, size_t> ext_allocations;
    // Batches created through dynaverse_hip_ext.h (freed with the owner).
    std::vector<BatchObj*> batches;
    // KZG SRS objects created through dynaverse_kzg.h (freed with the owner).
    std::vector<KzgSrs*> kzg;
    // Subproduct trees created through dynaverse_poly.h (freed with the owner).
    std::vector<PolyTree*> poly_trees;
    // Prepared G2 lines and the G1 doubling table of the pairing checks.
    PairingCache* pairing = nullptr;
    // Measured window width per log2 MSM size (dvgpu_ext_set_tuning); the
    // NTT stages per pass live in `domain`, the task shape in msm_config.
    WindowTable window_table;
    int msm_finalize_min = 0;  // 0 = built-in threshold for device-side window sums
    // Phase events of the phase-timed device MSM, made on its first call
    // (created together or not at all) and released with the device state.
    hipEvent_t msm_phase[MSM_PHASE_EVENTS] = {};
    hipEvent_t* msm_phase_events();

    Owner();
    ~Owner();
    void retire() noexcept;
    void release_device() noexcept;
};

// Helpers shared by the ABI translation units.
int fail(int status, const char* message, char* error, size_t capacity);

template <class Body>
int guarded(Owner* owner, char* error, size_t capacity, Body&& body) {
    if (!owner || owner->retired)
        return fail(-4, "backend retired after a device error; create a new backend", error, capacity);
    try {
        return body();
    } catch (const OwnerRetired&) {
        return fail(-4, "backend retired after a device error; create a new backend", error, capacity);
    } catch (const InvalidArgument& e) {
        return fail(-1, e.what(), error, capacity);
    } catch (const std::exception& e) {
        owner->retire();
        return fail(-2, e.what(), error, capacity);
    } catch (...) {
        owner->retire();
        return fail(-3, "unknown native exception", error, capacity);
    }
}

inline uint64_t now_ns() {
    return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                        std::chrono::steady_clock::now().time_since_epoch())
                        .count());
}

// Resident engine hooks (resident.hip).
void resident_destroy(ResidentState* state) noexcept;
DvgpuResidentStats resident_stats(const ResidentState* state);
DvgpuResidentDiagnostics resident_diagnostics(const ResidentState* state);
// Batch hooks (batch.hip).
void batch_destroy(BatchObj* b) noexcept;
// Plan the non-empty items of a batched MSM call; map[i] is the caller's
// index of planned MSM i (descriptors carry it as out_index).
MsmBatchPlan plan_items(const DvgpuMsmItem* items, size_t count, const Owner* o, std::vector<size_t>& map, bool g2);
// Setup hooks (setup.hip).
void setup_destroy(SetupState* state) noexcept;
// KZG hooks (kzg.hip).
void kzg_destroy(KzgSrs* srs) noexcept;
// Polynomial hooks (poly.hip).
void poly_tree_destroy(PolyTree* tree) noexcept;
// Pairing hooks (pairing.hip).
void pairing_cache_destroy(PairingCache* cache) noexcept;

}  // namespace dvgpu
