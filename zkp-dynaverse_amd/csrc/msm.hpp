// Variable-base multi-scalar multiplication on BN254 G1 / G2 (Pippenger),
// batched: one fixed sequence of kernels computes any number of MSMs of
// uneven sizes, each with its own window width.
#pragma once
#include <vector>

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
    int window_bits = 0;  // c for every MSM of a call; 0 selects by size
    int task_points = 16; // maximum points one accumulation thread adds (measured best on MI355X)
    int chunk = 8;        // buckets per first-level reduction thread
};

// Window width by log2 size (index 0..24), filled from measurement; an
// entry of 0 falls back to the built-in rule.
struct WindowTable {
    int c[25] = {0};
    int chunk[25] = {0};  // buckets per reduction thread by log2 size (0 = the config's)
};
int msm_window_for(size_t n, const WindowTable* table);

// One MSM of a batch, as the caller describes it.
struct MsmItem {
    size_t n;
    uint32_t point_base;   // first point (or first entry of the index array)
    uint32_t scalar_base;  // first scalar
    int c;                 // 0 = by size
    // Window-table mode (0 = off): the points are rows of `table_windows`
    // precomputed multiples [2^(c w)] B, so every digit of every window lands
    // in one bucket set and no doublings combine windows.  Point indices then
    // name rows; row r, window w is point r * table_windows + w.
    uint32_t table_windows = 0;
};

// Per-MSM descriptor as the kernels read it (all bases are exclusive
// prefix sums over the batch).
constexpr uint32_t MSM_NO_SLOT = 0xffffffffu;

struct MsmDesc {
    uint32_t n, c, windows, buckets;  // c: the widest window's bits; buckets 2^(c-1) per window
    uint32_t point_base, scalar_base;
    uint32_t thread_base, entry_base, bucket_base, chunk_base, window_base;
    uint32_t out_index;    // position of this MSM's result in the caller's list
    uint32_t result_slot;  // device finalize: also store the result at this point index
    uint32_t digit_windows;  // digits per scalar (= windows, or table_windows in table mode)
    uint32_t table;          // 1 in window-table mode (then windows == 1)
    uint32_t chunks;           // chunks of 2^chunk_shift points (histogram blocks per window; per MSM in table mode)
    uint32_t hist_block_base;  // first histogram block (chunks x digit_windows per MSM)
    uint32_t hist_base;        // first counter of its histograms (one bucket set per block)
    uint32_t chunk_shift;      // log2 of the points per histogram block
    uint32_t chunk;            // buckets per first-level reduction thread
    uint32_t plane_base;       // first bit-plane sum of this MSM (plane mode)    // Windows below `wide` are c bits wide, the others c - 1 (the widths sum
    // to at least 255, and the top window is not left with a few bits and
    // as few buckets for all points).
    uint32_t wide;
};

// Digits are ranked per block of 2^chunk_shift points and one window in LDS
// when every bucket set of a batch has at most MSM_LDS_BUCKETS buckets.  A
// block's histogram costs one counter per bucket, so blocks hold at least
// twice as many points as buckets (and 4096..32768 points).
constexpr uint32_t MSM_LDS_BUCKETS = 16384;
constexpr uint32_t MSM_DIGIT_THREADS = 1024;

struct MsmBatchPlan {
    std::vector<MsmDesc> descs;
    uint32_t total_threads = 0, total_entries = 0, total_buckets = 0, total_chunks = 0, total_windows = 0;
    uint32_t total_hist_blocks = 0, total_hist = 0, digit_threads = MSM_DIGIT_THREADS, max_window_chunks = 0;
    bool lds_digits = false;  // every bucket set fits the LDS histogram
    // Group placement (lds_digits batches without window tables whose bucket
    // sets are multiples of 64 and whose tasks hold at most 32 points): the
    // entries are staged per group of 64 buckets in the partial-sum buffer,
    // then placed group by group.
    bool group_place = false;
    size_t max_tasks = 0, max_heavy = 0;
    uint32_t task_points = 16;
    // Indexed points at or beyond this limit are counted in the workspace's
    // error word (total[3]) and read as point 0; only the finalize entry
    // point, which knows the point count, sets it.
    uint32_t point_limit = 0xffffffffu;
    // Plane mode (single MSMs whose windows are combined on the host): the
    // bucket reduction leaves, per window and bit k of the bucket weight, the
    // sum of the buckets whose weight has bit k set (total_planes values, each
    // summed in plane_slices parts first) instead of one value per window.
    // Chunk-plane mode (the same callers, larger bucket sets): the chunk
    // running sums keep their weighted and plain sums and the reduction
    // leaves, per window, the sum of the weighted ones and the bit-plane sums
    // of the plain ones over the chunk index (total_planes values).
    bool planes = false, chunk_planes = false;
    uint32_t total_planes = 0, plane_slices = 1;
};

MsmBatchPlan msm_batch_plan(const std::vector<MsmItem>& items, const MsmConfig& cfg,
                            const WindowTable* table = nullptr);

constexpr int MSM_PHASE_EVENTS = 8;

// Scratch owned by one backend; grows to the largest batch seen.
struct MsmWorkspace {
    DeviceBuffer digits, ranks, counts, offsets, sorted, tasks, partials, chunks, block_sums, total,
        points, scalars, windows, heavy, descs, finals, hist, order;
    HostBuffer host_windows, host_descs;
    // The descriptors last copied to `descs` (and that buffer's generation
    // then): a plan whose descriptors are already there skips the upload.
    std::vector<MsmDesc> uploaded;
    uint64_t uploaded_generation = ~uint64_t(0);
    // Optional phase events (MSM_PHASE_EVENTS of them) recorded by the
    // enqueue: start, digits, scan, scatter, tasks, accumulate, reduce, tree.
    hipEvent_t* phase_events = nullptr;
    DeviceBuffer* all[18] = {&digits, &ranks, &counts, &offsets, &sorted, &tasks,  &partials,
                             &chunks, &block_sums, &total, &points, &scalars, &windows, &heavy,
                             &descs,  &finals, &hist, &order};
    void attach(Quota* q) {
        for (DeviceBuffer* b : all) b->quota = q;
    }
    void release() {
        for (DeviceBuffer* b : all) b->release();
        host_windows.release();
        host_descs.release();
        uploaded.clear();
    }
};

// Size every buffer of `plan` and copy its descriptors to the device
// (enqueued on `stream`).  Must precede msm_batch_enqueue; a captured graph
// calls it once before capture.
void msm_batch_prepare(const MsmBatchPlan& plan, int xyzz_words, MsmWorkspace& ws, hipStream_t stream);

// points: device Montgomery affine words; point_index: optional indices
// into `points` (sparse MSMs over resident bases); scalars: canonical words.
// Leaves plan.total_windows XYZZ values in window_out (plan.total_planes
// bit-plane sums in plane mode).  Launches only
// kernels (no memset, no allocation), so it can be captured in a hipGraph.
template <class F>
void msm_batch_enqueue(const uint32_t* points, const uint32_t* point_index, const uint32_t* scalars,
                       const MsmBatchPlan& plan, MsmWorkspace& ws, hipStream_t stream, uint32_t* window_out);

// Host epilogue: for every MSM, sum_w 2^(w c) V_w then canonical affine.
template <class F>
void msm_batch_combine(const uint32_t* window_words, const MsmBatchPlan& plan, Affine<F>* out);

// Device epilogue, one lane per MSM: the same sum and affine conversion,
// written as canonical bytes to out + desc.out_index * CANON_BYTES and, where
// desc.result_slot != MSM_NO_SLOT and slot_points is given, as a Montgomery
// affine point at that index of slot_points.  Kernels only (capturable).
template <class F>
void msm_batch_finalize_enqueue(const uint32_t* window_words, const MsmBatchPlan& plan, const MsmWorkspace& ws,
                                uint8_t* out, uint32_t* slot_points, hipStream_t stream);

// Point the workspace's buffers at slices of `capacity` bytes of device
// memory at `base` (nothing is allocated; the window values go to
// ws.windows); returns the bytes used, or throws InvalidArgument when the
// plan does not fit.  Descriptors are not copied.
// Device bytes msm_run holds in its workspace after an MSM of n points (the
// single-MSM plan's buffers, each with its growth headroom).
uint64_t msm_run_workspace_bytes(size_t n, const MsmConfig& cfg, bool g2);

size_t msm_workspace_borrow(const MsmBatchPlan& plan, int xyzz_words, uint8_t* base, size_t capacity,
                            MsmWorkspace& ws);

// Zero `words` 32-bit words of device memory with a kernel (capturable).
void device_fill_words(uint32_t* p, size_t words, uint32_t value, hipStream_t stream);

// ---- single MSM conveniences ------------------------------------------------

struct MsmPlan {
    int c, windows;
    int wide = 0;  // windows below this index are c bits wide, the others c - 1
    uint32_t buckets;
    size_t n;
    int planes = 0;      // plane sums per window (the single-MSM path's window output)
    int chunk = 0;       // chunk-plane mode: buckets per chunk (planes = [U, Q_0, Q_1, ...]); 0: bucket planes
    size_t outputs = 0;  // XYZZ values msm_enqueue leaves in window_out: windows * planes
};
int msm_default_window(size_t n);
MsmPlan msm_plan(size_t n, const MsmConfig& cfg);

// One MSM; window_out holds room for `out_capacity` XYZZ values (at least
// msm_plan(n, cfg).outputs) and receives the returned plan's outputs.
template <class F>
MsmPlan msm_enqueue(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                    hipStream_t stream, const MsmConfig& cfg, uint32_t* window_out, size_t out_capacity);

template <class F>
Affine<F> msm_combine_windows(const uint32_t* window_words, const MsmPlan& plan);

template <class F>
Affine<F> msm_run(const uint32_t* points, const uint32_t* scalars, size_t n, MsmWorkspace& ws,
                  hipStream_t stream, const MsmConfig& cfg);

// Canonical little-endian points (64 B G1 / 128 B G2, all-zero = infinity)
// on the device -> Montgomery words (all-zero stays all-zero).
template <class F>
void points_to_mont(const uint8_t* canonical, uint32_t* mont, size_t n, hipStream_t stream);

// Canonical affine -> 64/128 output bytes.
template <class F>
void affine_to_bytes(const Affine<F>& a, uint8_t* out);

}  // namespace dvgpu
