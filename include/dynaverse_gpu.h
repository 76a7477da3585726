/*
 * dynaverse_gpu.h -- C ABI of the MI355X (gfx950) BN254 proving library.
 *
 * Drop-in for the ZKP-Dynaverse native bridge: every symbol, constant,
 * struct layout and status code below is the one the reference's Rust
 * wrappers bind (crates/gpu/src/icicle.rs:35-57, resident.rs:81-92,
 * resident_profile.rs, setup.rs:47-52, g2_validation.rs:84-95).  The
 * declarations are restated from native/include/dynaverse_gpu.h:1-194 and
 * the extra definitions in native/src/icicle_msm.cu:474,494,578,750 and
 * native/include/{dynaverse_g2_experiments.h,icicle_transform_packet.cuh,
 * dynaverse_bn254_{tower,lines,miller}.h}.
 *
 * Contract (unchanged from the reference):
 *  - Status 0 = success; -1 invalid argument (owner stays usable); -2 device
 *    execution error (owner retired); -3 unknown failure (owner retired);
 *    -4 owner already retired; -5 G2 content rejected (report published,
 *    output untouched).  Messages are NUL-terminated into (error, capacity).
 *  - All host buffers are borrowed for the duration of the call only; calls
 *    are synchronous and outputs are written only on success.
 *  - One caller thread per owner; distinct owners may run concurrently.
 *  - Field elements: 32-byte little-endian canonical integers.  G1 affine:
 *    64 bytes x||y; G2 affine: 128 bytes x.c0||x.c1||y.c0||y.c1; infinity is
 *    all zero bytes.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

/* Size limits (reference dynaverse_gpu.h:4-10). */
enum { DVGPU_MAX_MSM_SIZE = 1 << 20 };
enum { DVGPU_MAX_NTT_SIZE = 1 << 20 };
enum { DVGPU_MAX_TRANSFORM_MATRIX_SIZE = 1 << 21 };
enum { DVGPU_MAX_MSM_BATCH_SIZE = 64 };
enum { DVGPU_MAX_LARGE_BUCKET_FACTOR = 100 };

#ifdef __cplusplus
extern "C" {
#endif

/* ---- owner lifecycle (reference dynaverse_gpu.h:18-27) ------------------ */
/* Creates an owner bound to HIP device 0 with its own stream and scratch. */
int dvgpu_backend_create(void** backend, char* error, size_t error_capacity);
/* factor 0..100 is accepted and recorded; this library's bucket scheduler
 * already splits heavy buckets, so the factor selects nothing else. */
int dvgpu_backend_create_with_large_bucket_factor(uint32_t factor, void** backend,
                                                  char* error, size_t error_capacity);
void dvgpu_backend_destroy(void* backend);
int dvgpu_backend_device_identity(void* backend, char* output, size_t output_capacity,
                                  char* error, size_t error_capacity);

/* ---- MSM / NTT (reference dynaverse_gpu.h:33-52) ------------------------ */
int dvgpu_bn254_msm(void* backend, const uint8_t* points, const uint8_t* scalars, size_t length,
                    uint8_t* result, char* error, size_t error_capacity);
int dvgpu_bn254_g2_msm(void* backend, const uint8_t* points, const uint8_t* scalars,
                       size_t length, uint8_t* result, char* error, size_t error_capacity);
/* Natural-order Fr transform; inverse includes 1/length.  root: canonical
 * primitive 2^20-th root (order checked).  length: power of two <= 2^20. */
int dvgpu_bn254_ntt(void* backend, const uint8_t* values, size_t length, int inverse,
                    const uint8_t* root, uint8_t* result, char* error, size_t error_capacity);
/* count contiguous MSMs of `length` points each (zero padded); count <= 64,
 * length * count <= 2^20; results count * 64 bytes. */
int dvgpu_bn254_msm_batch(void* backend, const uint8_t* points, const uint8_t* scalars,
                          size_t length, size_t count, uint8_t* results,
                          char* error, size_t error_capacity);

typedef struct DvgpuBatchStats {
    uint64_t icicle_calls;   /* native batch calls attempted */
    uint64_t buffer_grows;   /* persistent input growth events */
    uint64_t max_batch_size;
    uint64_t input_capacity; /* retained point capacity */
} DvgpuBatchStats;

/* Fixed-base setup ablation modes 0/1/2 and stage timings
 * (reference dynaverse_gpu.h:59-66). */
typedef struct DvgpuSetupStats {
    uint64_t table_builds[2], calls[2], points[2], table_bytes;
    uint64_t device_ns[5], native_wall_ns; /* table, H2D, generate, normalize, D2H */
} DvgpuSetupStats;
int dvgpu_setup_mode(void* backend, int mode);
int dvgpu_setup_stats(void* backend, DvgpuSetupStats* stats);

/* ---- G2 subgroup validation (reference dynaverse_gpu.h:67-89) ---------- */
typedef struct DvgpuG2ValidationReport {
    uint64_t checked_points, invalid_points, first_invalid;
    uint32_t failure_mask, algorithm;
} DvgpuG2ValidationReport;
typedef struct DvgpuG2ValidationStats {
    uint64_t parse_curve_device_ns, subgroup_device_ns, reduce_device_ns;
    uint64_t h2d_device_ns, d2h_device_ns, native_wall_ns;
    uint64_t checked_points, accepted_points;
} DvgpuG2ValidationStats;
/* Masks per point: 1 encoding, 2 curve, 4 subgroup.  n <= 4096; algorithm
 * 1 = [r]P, 2 = psi(P) == [6x^2]P, 3 = 63-bit seed relation. */
int dvgpu_bn254_g2_validation_inspect(void* owner, const uint8_t* wire, size_t n,
                                      uint32_t algorithm, uint32_t* point_masks,
                                      DvgpuG2ValidationReport* report,
                                      char* error, size_t error_capacity);
int dvgpu_g2_validation_stats(void* owner, DvgpuG2ValidationStats* stats);
typedef struct DvgpuG2ValidationTraffic {
    uint64_t scalar_h2d_bytes, point_h2d_bytes, point_d2h_bytes, status_d2h_bytes;
} DvgpuG2ValidationTraffic;
int dvgpu_g2_validation_traffic(void* owner, DvgpuG2ValidationTraffic* stats);
/* [s_i] g2 for n <= 4096 scalars, validated on the device before publication;
 * -5 publishes only the rejection report. */
int dvgpu_bn254_fixed_base_g2_checked(void* owner, const uint8_t* scalars, size_t n,
                                      uint32_t algorithm, uint8_t* output,
                                      DvgpuG2ValidationReport* report,
                                      char* error, size_t error_capacity);

int dvgpu_backend_batch_stats(void* backend, DvgpuBatchStats* stats);
typedef struct DvgpuMemoryInfo { uint64_t free_bytes; uint64_t total_bytes; } DvgpuMemoryInfo;
int dvgpu_backend_memory_info(void* backend, DvgpuMemoryInfo* info,
                              char* error, size_t error_capacity);

/* ---- resident operator engine (reference dynaverse_gpu.h:99-158) ------- */
enum DvgpuResidentOpcode {
    DVGPU_RESIDENT_ALLOC = 1, DVGPU_RESIDENT_LENGTH, DVGPU_RESIDENT_KIND,
    DVGPU_RESIDENT_BEGIN, DVGPU_RESIDENT_ROLLBACK, DVGPU_RESIDENT_WRITE,
    DVGPU_RESIDENT_READ, DVGPU_RESIDENT_COPY, DVGPU_RESIDENT_REVERSE,
    DVGPU_RESIDENT_TRANSPOSE, DVGPU_RESIDENT_TRANSFORM, DVGPU_RESIDENT_ZIP,
    DVGPU_RESIDENT_SCALE, DVGPU_RESIDENT_MSM_G1, DVGPU_RESIDENT_MSM_G2,
    DVGPU_RESIDENT_ALLOC_BASES
};
/* Reference icicle_transform_packet.cuh:7: a = input, b = scratch, c = output,
 * length = full FFT size (<= 4096), flags = inverse. */
enum { DVGPU_RESIDENT_TRANSFORM_PACKET = 17, DVGPU_MAX_TRANSFORM_PACKET_SIZE = 4096 };
enum { DVGPU_READ_DEVICE_AFFINE = 1, DVGPU_READ_DEVICE_AFFINE_MAX = 256 };
typedef struct DvgpuResidentOp {
    uint32_t opcode, kind, flags, reserved;
    uint64_t a, b, c;
    uint64_t a_offset, b_offset, c_offset, length, rows, cols;
    const uint8_t* input;
    uint8_t* output;
} DvgpuResidentOp;
typedef struct DvgpuResidentStats {
    uint64_t h2d_bytes, d2h_bytes, bytes_allocated;
    uint64_t fr_transform_elements, g1_transform_elements;
    uint64_t msm_g1_points, msm_g2_points, calls;
} DvgpuResidentStats;
typedef struct DvgpuResidentDiagnostics {
    uint64_t resident_allocations, temporary_allocations, temporary_peak_bytes;
    uint64_t journal_saved_bytes, journal_saved_ranges, journal_peak_bytes;
    uint64_t device_operation_ns, device_operations;
    uint64_t physical_h2d_bytes, physical_d2h_bytes, explicit_d2d_bytes;
    uint64_t journal_metadata_capacity_bytes;
} DvgpuResidentDiagnostics;
int dvgpu_resident_configure(void* backend, uint64_t memory_limit_bytes,
                             uint64_t journal_bytes, const uint8_t* root,
                             char* error, size_t error_capacity);
int dvgpu_resident_call(void* backend, const DvgpuResidentOp* operation,
                        uint64_t* result, char* error, size_t error_capacity);
void dvgpu_resident_seal(void* backend);
void dvgpu_resident_commit(void* backend);
int dvgpu_resident_stats(void* backend, DvgpuResidentStats* stats);
int dvgpu_resident_diagnostics(void* backend, DvgpuResidentDiagnostics* stats);
typedef struct DvgpuResidentProfile {
    uint64_t native_wall_ns, journal_host_ns, conversion_host_ns;
    uint64_t operation_host_ns, wait_host_ns;
    uint64_t event_total_ns, event_journal_ns, event_operation_ns, event_mask;
    uint64_t h2d_bytes, d2h_bytes, d2d_bytes;
} DvgpuResidentProfile;
int dvgpu_resident_call_profiled(void* backend, const DvgpuResidentOp* operation,
                                 uint64_t* result, DvgpuResidentProfile* profile,
                                 char* error, size_t error_capacity);

/* ---- MSM diagnostics (reference dynaverse_gpu.h:161-191) --------------- */
typedef struct DvgpuMsmTimings {
    uint64_t conversion_ns; /* host input handling */
    uint64_t icicle_ns;     /* device phase: upload, kernels, synchronisation */
    uint64_t output_ns;     /* host window combination and encoding */
} DvgpuMsmTimings;
int dvgpu_bn254_msm_profiled(void* backend, const uint8_t* points, const uint8_t* scalars,
                             size_t length, uint8_t* result, DvgpuMsmTimings* timings,
                             char* error, size_t error_capacity);
enum { DVGPU_TIMELINE_PHASES = 9, DVGPU_TIMELINE_KERNELS = 10 };
typedef struct DvgpuMsmTimeline {
    uint64_t wall_ns[DVGPU_TIMELINE_PHASES];
    uint64_t event_ns[DVGPU_TIMELINE_PHASES];
    uint64_t calls[DVGPU_TIMELINE_PHASES];
    uint64_t event_calls[DVGPU_TIMELINE_PHASES];
    uint64_t bytes[DVGPU_TIMELINE_PHASES];
    uint64_t kernel_ns[DVGPU_TIMELINE_KERNELS];
    uint64_t kernel_calls[DVGPU_TIMELINE_KERNELS];
    uint64_t instrumentation_inclusive_ns;
    uint64_t instrumentation_total_ns;
    uint64_t c;
    uint64_t bitsize;
    uint64_t windows;
    uint64_t buckets;
} DvgpuMsmTimeline;
int dvgpu_bn254_msm_timeline(void* backend, const uint8_t* points, const uint8_t* scalars,
                             size_t length, uint8_t* result, const char* label,
                             DvgpuMsmTimings* timings, DvgpuMsmTimeline* timeline,
                             char* error, size_t error_capacity);

/* ---- symbols the Rust side declares outside the reference header ------- */
/* Reference icicle_msm.cu:474 -- pre-grow the batch input buffers. */
int dvgpu_reserve_batch(void* owner, size_t points, char* error, size_t error_capacity);
/* Reference icicle_msm.cu:494 -- [s_i] G for G1 (g2 = 0) or G2 (g2 = 1),
 * n <= 4096 canonical scalars, canonical affine output. */
int dvgpu_bn254_fixed_base(void* owner, const uint8_t* scalars, size_t length, int g2,
                           uint8_t* output, char* error, size_t error_capacity);
/* Reference icicle_msm.cu:750 -- n <= 129 homogeneous (X, Y, Z) G2 points of
 * 192 bytes; writes psi, psi^2, psi^3 of each (576 bytes per point). */
int dvgpu_g2_psi_probe(void* owner, const uint8_t* points, size_t n, uint8_t* output,
                       char* error, size_t error_capacity);
/* Reference icicle_msm.cu:578 -- test-only: flip one bit of output point
 * `index` of the next checked G2 chunk. */
int dvgpu_g2_test_corrupt_next(void* owner, uint64_t index);

/* G2 kernel configuration (reference dynaverse_g2_experiments.h:6-17). */
typedef struct DvgpuG2KernelConfig { uint32_t threads, schedule, layout, fused; } DvgpuG2KernelConfig;
typedef struct DvgpuG2ExperimentStats {
    uint64_t repack_device_ns, subgroup_kernel_ns, logical_repack_bytes;
} DvgpuG2ExperimentStats;
typedef struct DvgpuG2KernelResources {
    uint32_t registers, block_threads, active_blocks, max_threads_per_sm;
    uint64_t local_bytes, shared_bytes;
    uint32_t normalize_registers, normalize_active_blocks;
    uint64_t normalize_local_bytes, normalize_shared_bytes;
} DvgpuG2KernelResources;
int dvgpu_g2_kernel_config_set(void* owner, const DvgpuG2KernelConfig* config);
int dvgpu_g2_kernel_config_get(void* owner, DvgpuG2KernelConfig* config);
int dvgpu_g2_kernel_resources(void* owner, uint32_t algorithm, DvgpuG2KernelResources* output);
int dvgpu_g2_experiment_stats(void* owner, DvgpuG2ExperimentStats* output);

/* Pairing-tower diagnostics (reference dynaverse_bn254_{tower,lines,miller}.h).
 * Not on the proving path: present so the Rust probe modules link; they
 * validate their arguments and return -1 with an explanatory message. */
enum { DVGPU_TOWER_MAX_ITEMS = 64 };
enum { DVGPU_TOWER_FQ2 = 2, DVGPU_TOWER_FQ6 = 6, DVGPU_TOWER_FQ12 = 12 };
enum { DVGPU_TOWER_ADD = 1, DVGPU_TOWER_SUBTRACT = 2, DVGPU_TOWER_MULTIPLY = 3,
       DVGPU_TOWER_SQUARE = 4, DVGPU_TOWER_INVERSE = 5, DVGPU_TOWER_FROBENIUS = 6 };
int dvgpu_bn254_tower_inspect(void* owner, uint32_t degree, uint32_t operation, uint32_t power,
                              const uint8_t* left, size_t left_bytes, const uint8_t* right,
                              size_t right_bytes, uint8_t* output, size_t output_bytes,
                              uint32_t* statuses, size_t statuses_count,
                              char* error, size_t error_capacity);
int dvgpu_bn254_lines_inspect(void* owner, const uint8_t* points, size_t points_bytes,
                              uint8_t* output, size_t output_bytes, uint32_t* counts,
                              size_t counts_count, uint32_t* statuses, size_t statuses_count,
                              char* error, size_t error_capacity);
int dvgpu_bn254_miller_inspect(void* owner, const uint8_t* p, size_t p_bytes, const uint8_t* q,
                               size_t q_bytes, uint8_t* output, size_t output_bytes,
                               uint32_t* metadata, size_t metadata_count,
                               char* error, size_t error_capacity);

#ifdef __cplusplus
}
#endif
