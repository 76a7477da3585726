/*
 * dynaverse_hip_ext.h -- device-resident entry points of the MI355X library.
 *
 * These have no counterpart in the reference ABI.  They run the same kernels
 * as dynaverse_gpu.h on buffers that already live in device memory (HBM), so
 * that callers which keep data resident (the batching and serving layers,
 * and throughput measurement) skip the host<->device copies.  Pointers are
 * device addresses on HIP device 0; every call is synchronous on the owner's
 * stream and returns the dynaverse_gpu.h status codes.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "dynaverse_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device memory owned by this owner (freed by dvgpu_ext_free or when the
 * owner is destroyed or retired).  Copies are synchronous. */
int dvgpu_ext_alloc(void* owner, size_t bytes, void** ptr, char* error, size_t error_capacity);
int dvgpu_ext_free(void* owner, void* ptr);
int dvgpu_ext_copy_to_device(void* owner, void* dst, const void* src, size_t bytes,
                             char* error, size_t error_capacity);
int dvgpu_ext_copy_to_host(void* owner, void* dst, const void* src, size_t bytes,
                           char* error, size_t error_capacity);

/* Device-to-device copy by a plain load/store kernel (16-byte vectors):
 * the reference copy for HBM traffic-counter calibration.  bytes % 16 == 0. */
int dvgpu_ext_device_copy(void* owner, void* dst, const void* src, size_t bytes,
                          char* error, size_t error_capacity);

/* The owner's HIP stream (hipStream_t), for callers that time with events. */
int dvgpu_ext_stream(void* owner, void** stream);

/* Canonical little-endian affine points (64 B G1 / 128 B G2, zero = infinity)
 * -> Montgomery words, device to device. g2 selects the group. */
int dvgpu_ext_points_to_mont(void* owner, const void* canonical, void* mont, size_t n, int g2,
                             char* error, size_t error_capacity);

/* MSM over device-resident Montgomery points and canonical scalars.
 * result: 64 (G1) or 128 (G2) canonical bytes on the host.  window_bits 0
 * picks the default.  phase_ms, when not null, receives MSM_PHASE_EVENTS-1 = 7
 * event-timed phase durations: digits, scan, scatter, tasks, accumulate,
 * bucket reduce, window tree. */
int dvgpu_ext_msm_device(void* owner, const void* points_mont, const void* scalars, size_t n,
                         int g2, int window_bits, uint8_t* result, float* phase_ms,
                         char* error, size_t error_capacity);

/* rows x 2^log_n natural-order NTTs on canonical device words; in may equal
 * out.  root as in dvgpu_bn254_ntt.  kernel_ms (optional) receives the
 * event-timed duration of the transform. */
int dvgpu_ext_ntt_device(void* owner, const void* in, void* out, size_t length, size_t rows,
                         int inverse, const uint8_t* root, float* kernel_ms,
                         char* error, size_t error_capacity);

/* [s_i] G for n device-resident canonical scalars into device Montgomery
 * affine points (64 / 128 bytes each): device-side SRS generation. */
int dvgpu_ext_fixed_base_device(void* owner, const void* scalars, size_t n, int g2, void* points_mont,
                                char* error, size_t error_capacity);

/* ---- batching: many uneven MSMs and NTT rows in one launch sequence ---- */

/* Launch shapes chosen per size, e.g. by an autotuning sweep.  Every entry
 * of 0 keeps the built-in choice.  Applies to every later call of the owner
 * (plain, extension and batch entry points) and to batches created later. */
typedef struct DvgpuTuning {
    int32_t msm_window_bits[25];  /* by log2 MSM size: window width c, 4..20 */
    int32_t ntt_pass_stages[21];  /* by log2 NTT length: butterfly stages per
                                     LDS round trip, 1 (radix-2) or 2..11
                                     (radix-4) */
    int32_t msm_task_points;      /* max points one accumulation task adds, 1..1024 */
    int32_t msm_chunk;            /* buckets per first-level reduction thread: 1, 2, 4, 8 */
    int32_t msm_device_finalize_min; /* batches of at least this many MSMs sum their
                                        windows on the device (one lane per MSM) */
    int32_t msm_chunk_by_log2[25];  /* buckets per reduction thread by log2 MSM size
                                       (0 = msm_chunk) */
} DvgpuTuning;
int dvgpu_ext_set_tuning(void* owner, const DvgpuTuning* tuning);
/* The window width the owner uses for an MSM of n points (positive), or a
 * negative status. */
int dvgpu_ext_msm_window(void* owner, uint64_t n);
int dvgpu_ext_get_tuning(void* owner, DvgpuTuning* tuning);

typedef struct DvgpuMsmItem {
    uint64_t n;              /* points in this MSM */
    uint32_t point_base;     /* first point, or first entry of point_index */
    uint32_t scalar_base;    /* first scalar */
    int32_t window_bits;     /* 0 = from the tuning table */
    int32_t reserved;
} DvgpuMsmItem;

/* count MSMs over device-resident Montgomery points (G1 or G2) and canonical
 * scalars in one kernel sequence; point_index (optional, device u32) selects
 * points for sparse MSMs.  results: count x 64 (G1) / 128 (G2) host bytes.
 * Ranges are the caller's to keep (raw device pointers carry no length):
 * every item's points, indices and scalars must lie inside their buffers;
 * dvgpu_ext_msm_batch_finalize takes the lengths and checks them. */
int dvgpu_ext_msm_batch_device(void* owner, const void* points, const void* point_index,
                               const void* scalars, const DvgpuMsmItem* items, size_t count, int g2,
                               uint8_t* results, float* phase_ms, char* error, size_t error_capacity);

/* Same MSMs with the window sums finished on the device (one lane per MSM):
 * results_device (device, count x 64 / 128 bytes; null = internal) receives
 * the canonical affine results, results_host (optional) a copy.  With
 * slot_points non-null, MSM i also stores its result as a Montgomery affine
 * point at index result_slots[i] (host array, UINT32_MAX = none) of
 * slot_points, after every MSM of the call has read its inputs: a result may
 * replace one of its own input points (a cached commitment updated in place).
 * Bounds: item ranges are checked against point_count / index_count /
 * scalar_count and slots against slot_count on the host; indexed points at or
 * beyond point_count are caught on the device, and then the call fails with
 * -1 and stores no slot.  Empty MSMs give the identity and may not name a
 * slot. */
int dvgpu_ext_msm_batch_finalize(void* owner, const void* points, size_t point_count,
                                 const void* point_index, size_t index_count,
                                 const void* scalars, size_t scalar_count,
                                 const DvgpuMsmItem* items, size_t count, int g2,
                                 void* results_device, uint8_t* results_host,
                                 const uint32_t* result_slots, void* slot_points, size_t slot_count,
                                 char* error, size_t error_capacity);

/* Bytes of the device array behind dvgpu_ext_msm_batch_finalize_noalloc
 * (one per device the library runs on, held until the process ends). */
#define DVGPU_NOALLOC_SCRATCH_BYTES (64u << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The dvgpu_ext_msm_batch_finalize computation (no result
 * slots) with every intermediate buffer taken from a fixed device array of
 * DVGPU_NOALLOC_SCRATCH_BYTES: the call makes no hipMalloc, hipFree or other
 * allocation.  results_device (device, count x 64 / 128 bytes) is required;
 * results_host (optional, host) receives a copy.  A batch whose buffers do not
 * fit the array fails with -1 and leaves the owner usable: at the default
 * window widths one G1 MSM of up to 2^17 points or one G2 MSM of 2^16 fits,
 * a G1 MSM of 2^18 does not.
 * Every call on one device shares the array: calls that overlap on one device
 * (two threads or two owners) race, and callers serialise them. */
int dvgpu_ext_msm_batch_finalize_noalloc(void* owner, const void* points, size_t point_count,
                                         const void* point_index, size_t index_count,
                                         const void* scalars, size_t scalar_count,
                                         const DvgpuMsmItem* items, size_t count, int g2,
                                         void* results_device, uint8_t* results_host,
                                         char* error, size_t error_capacity);

typedef struct DvgpuNttRows {
    const void* in;          /* device canonical words */
    void* out;
    uint64_t length, rows;   /* rows x length, length a power of two <= 2^20 */
    int32_t inverse, reserved;
} DvgpuNttRows;

/* All row batches in one launch sequence (one launch per pass for every
 * entry, at most two; the last pass stores in natural order), whatever their
 * lengths.  Canonical in and out; entries may not overlap each other (each
 * may run in place). */
int dvgpu_ext_ntt_batch_device(void* owner, const DvgpuNttRows* rows, size_t count, const uint8_t* root,
                               float* kernel_ms, char* error, size_t error_capacity);

typedef struct DvgpuBatchSpec {
    const void* g1_points; const void* g1_index; const void* g1_scalars;
    const DvgpuMsmItem* g1_items; uint64_t g1_count;
    const void* g2_points; const void* g2_index; const void* g2_scalars;
    const DvgpuMsmItem* g2_items; uint64_t g2_count;
    const DvgpuNttRows* ntt; uint64_t ntt_count;
    const void* g2_wire; uint64_t g2_wire_n;   /* device 128-byte wire points */
    uint32_t g2_algorithm, reserved;
    void* g2_masks;                            /* device u32[g2_wire_n] */
    const uint8_t* root;                       /* host, 2^20-th root for the NTTs */
} DvgpuBatchSpec;

/* Plan, size and warm a batch; with use_graph the whole sequence (G1 MSMs,
 * G2 MSMs, NTT rows, G2 validation, result copies) is captured into one
 * hipGraph that dvgpu_ext_batch_run replays with a single launch.  NTT
 * entries run in order; consecutive entries that do not overlap run fused
 * (an entry may read an earlier entry's output). */
int dvgpu_ext_batch_create(void* owner, const DvgpuBatchSpec* spec, int use_graph, void** batch,
                           char* error, size_t error_capacity);
/* Runs the batch once; outputs are host buffers (any may be null). */
int dvgpu_ext_batch_run(void* owner, void* batch, uint8_t* g1_results, uint8_t* g2_results,
                        DvgpuG2ValidationReport* report, float* device_ms,
                        char* error, size_t error_capacity);
int dvgpu_ext_batch_destroy(void* owner, void* batch);

/* ---- serving: batched sparse updates of device-resident session states ---- */

/* Window table of n resident bases for the serving path: row i receives the
 * ceil(255 / window_bits) points [s_i 2^(c w)] g1, w = 0.. (Montgomery affine,
 * 64 bytes each), from n device canonical scalars s_i. */
int dvgpu_ext_window_table_device(void* owner, const void* scalars, size_t n, int window_bits, void* rows,
                                  char* error, size_t error_capacity);

/* One sparse update of one session.  Rows refer to the `points` arena viewed
 * as rows of ceil(255 / window_bits) points; field indices to the `fr` arena
 * (canonical Fr, 32 bytes each).  A session of size n = 2^log_n holds the
 * window tables of its basis P (rows srs_row..+n) and opening basis U (rows
 * opening_row..+n), the cached commitment C and opening W as XYZZ points
 * (128 bytes at the start of their rows; all-zero = identity), coefficients
 * f, evaluations NTT(f), the linear-term vector l and the cached value
 * y = <l, f>. */
typedef struct DvgpuUpdate {
    uint32_t srs_row, opening_row, commitment_row, opening_cache_row;
    uint64_t coeff_offset, eval_offset, linear_offset, value_slot;
    uint32_t log_n;
    uint32_t k;               /* changed coefficients (0 allowed) */
    const uint32_t* indices;  /* host: k distinct indices < n */
    const uint8_t* deltas;    /* host: k canonical Fr */
} DvgpuUpdate;

/* Applies every update of the batch in one launch sequence:
 *   C += sum d_j P_{i_j},  W += sum d_j U_{i_j}   (in place, XYZZ),
 *   f_{i_j} += d_j,  y += sum d_j l_{i_j},  evaluations = NTT(f) (root given).
 * results: count x 160 host bytes = new C (64) || new W (64) || new y (32),
 * canonical.  Updates of one batch must not share session state (checked,
 * as are all ranges); device_ms (optional) is the kernel time. */
int dvgpu_ext_update_batch(void* owner, void* points, size_t point_count, int window_bits, void* fr,
                           size_t fr_count, const DvgpuUpdate* updates, size_t count, const uint8_t* root,
                           uint8_t* results, float* device_ms, char* error, size_t error_capacity);

/* Sizes every workspace of the owner for update batches of up to
 * max_updates updates with max_entries changed coefficients and
 * max_elements NTT elements in total, so that later batches within those
 * bounds allocate (and free) nothing. */
int dvgpu_ext_update_reserve(void* owner, size_t max_updates, size_t max_entries, size_t max_elements,
                             int window_bits, const uint8_t* root, char* error, size_t error_capacity);

#ifdef __cplusplus
}
#endif
