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

/* Window width per log2 MSM size (25 entries, 0 = built-in rule), e.g. the
 * table an autotuning sweep measured.  Used whenever window_bits is 0. */
int dvgpu_ext_set_window_table(void* owner, const int32_t* window_bits_by_log2);

typedef struct DvgpuMsmItem {
    uint64_t n;              /* points in this MSM */
    uint32_t point_base;     /* first point, or first entry of point_index */
    uint32_t scalar_base;    /* first scalar */
    int32_t window_bits;     /* 0 = from the window table */
    int32_t reserved;
} DvgpuMsmItem;

/* count MSMs over device-resident Montgomery points (G1 or G2) and canonical
 * scalars in one kernel sequence; point_index (optional, device u32) selects
 * points for sparse MSMs.  results: count x 64 (G1) / 128 (G2) host bytes. */
int dvgpu_ext_msm_batch_device(void* owner, const void* points, const void* point_index,
                               const void* scalars, const DvgpuMsmItem* items, size_t count, int g2,
                               uint8_t* results, float* phase_ms, char* error, size_t error_capacity);

typedef struct DvgpuNttRows {
    const void* in;          /* device canonical words */
    void* out;
    uint64_t length, rows;   /* rows x length, length a power of two <= 2^20 */
    int32_t inverse, reserved;
} DvgpuNttRows;

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
 * hipGraph that dvgpu_ext_batch_run replays with a single launch. */
int dvgpu_ext_batch_create(void* owner, const DvgpuBatchSpec* spec, int use_graph, void** batch,
                           char* error, size_t error_capacity);
/* Runs the batch once; outputs are host buffers (any may be null). */
int dvgpu_ext_batch_run(void* owner, void* batch, uint8_t* g1_results, uint8_t* g2_results,
                        DvgpuG2ValidationReport* report, float* device_ms,
                        char* error, size_t error_capacity);
int dvgpu_ext_batch_destroy(void* owner, void* batch);

#ifdef __cplusplus
}
#endif
