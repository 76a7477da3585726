// Variable-base multi-scalar multiplication on BN254 G1 / G2 (Pippenger).
#pragma once
#include "common.hpp"
#include "curve.hpp"

namespace dvgpu {

using bn254::Affine;
using bn254::Fq2Ops;
using bn254::FqOps;
using bn254::Xyzz;

// Memory layouts of one point, in 32-bit words.
template <class F>
struct PointLayout;
template <>
struct PointLayout<FqOps> {
    static constexpr int AFFINE_WORDS = 16;  // x || y
    static constexpr int XYZZ_WORDS = 32;    // X || Y || ZZ || ZZZ
    static constexpr int CANON_BYTES = 64;
};
template <>
struct PointLayout<Fq2Ops> {
    static constexpr int AFFINE_WORDS = 32;  // x.c0 || x.c1 || y.c0 || y.c1
    static constexpr int XYZZ_WORDS = 64;
    static constexpr int CANON_BYTES = 128;
};

struct MsmConfig {
    int window_bits = 0;  // c; 0 selects by size
    int task_points = 32; // maximum points one accumulation thread adds
    int chunk = 8;        // buckets per first-level reduction thread
    int tree_fanin = 4;   // values per thread in later reduction levels
};

// Scratch owned by one backend; grows to the largest MSM seen.
struct MsmWorkspace {
    DeviceBuffer digits, ranks, counts, offsets, sorted, tasks, partials, chunks, tree_a, tree_b,
        block_sums, total, points, scalars, windows, heavy;
    HostBuffer host_windows;
    // Optional phase events (MSM_PHASE_EVENTS of them) recorded by msm_enqueue:
    // start, digits, scan, scatter, tasks, accumulate, reduce, tree.
    hipEvent_t* phase_events = nullptr;
    void attach(Quota* q) {
        for (DeviceBuffer* b : {&digits, &ranks, &counts, &offsets, &sorted, &tasks, &partials,
                                &chunks, &tree_a, &tree_b, &block_sums, &total, &points, &scalars, &windows, &heavy})
            b->quota = q;
    }
    void release() {
        for (DeviceBuffer* b : {&digits, &ranks, &counts, &offsets, &sorted, &tasks, &partials,
                                &chunks, &tree_a, &tree_b, &block_sums, &total, &points, &scalars, &windows, &heavy})
            b->release();
        host_windows.release();
    }
};

constexpr int MSM_PHASE_EVENTS = 8;

int msm_default_window(size_t n);

// Plan of one MSM: window layout and buffer sizes.
struct MsmPlan {
    int c, windows;
    uint32_t buckets;      // 2^(c-1) per window
    size_t n;
    size_t max_tasks;
};
MsmPlan msm_plan(size_t n, const MsmConfig& cfg);

// points: device Montgomery affine words; scalars: device canonical words
// (8 per scalar).  Enqueues the device phases on `stream` and leaves one
// XYZZ value per window in ws.host_windows after an implicit synchronize,
// then combines the windows on the host.  Returns the canonical affine sum.
template <class F>
Affine<F> msm_run(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                  hipStream_t stream, const MsmConfig& cfg);

// Device-only part: leaves W window values (XYZZ words) in `window_out`.
template <class F>
MsmPlan msm_enqueue(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                    hipStream_t stream, const MsmConfig& cfg, uint32_t* window_out);

// Host epilogue: sum_w 2^(w c) V_w, then canonical affine.
template <class F>
Affine<F> msm_combine_windows(const uint32_t* window_words, const MsmPlan& plan);

// Canonical little-endian points (64 B G1 / 128 B G2, all-zero = infinity)
// on the device -> Montgomery words (all-zero stays all-zero).
template <class F>
void points_to_mont(const uint8_t* canonical, uint32_t* mont, size_t n, hipStream_t stream);

// Canonical affine -> 64/128 output bytes.
template <class F>
void affine_to_bytes(const Affine<F>& a, uint8_t* out);

}  // namespace dvgpu
