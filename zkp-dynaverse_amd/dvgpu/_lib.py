"""ctypes binding of libdynaverse_hip.so (the C ABI in include/*.h).

The library is loaded from this directory only.  A missing library is an
error: there is no CPU fallback anywhere in this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdynaverse_hip.so")

c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_char_p = ctypes.c_char_p


class DvgpuBatchStats(ctypes.Structure):
    _fields_ = [("icicle_calls", ctypes.c_uint64), ("buffer_grows", ctypes.c_uint64),
                ("max_batch_size", ctypes.c_uint64), ("input_capacity", ctypes.c_uint64)]


class DvgpuSetupStats(ctypes.Structure):
    _fields_ = [("table_builds", ctypes.c_uint64 * 2), ("calls", ctypes.c_uint64 * 2),
                ("points", ctypes.c_uint64 * 2), ("table_bytes", ctypes.c_uint64),
                ("device_ns", ctypes.c_uint64 * 5), ("native_wall_ns", ctypes.c_uint64)]


class DvgpuG2ValidationReport(ctypes.Structure):
    _fields_ = [("checked_points", ctypes.c_uint64), ("invalid_points", ctypes.c_uint64),
                ("first_invalid", ctypes.c_uint64), ("failure_mask", ctypes.c_uint32),
                ("algorithm", ctypes.c_uint32)]


class DvgpuG2ValidationStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "parse_curve_device_ns", "subgroup_device_ns", "reduce_device_ns", "h2d_device_ns",
        "d2h_device_ns", "native_wall_ns", "checked_points", "accepted_points")]


class DvgpuG2ValidationTraffic(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "scalar_h2d_bytes", "point_h2d_bytes", "point_d2h_bytes", "status_d2h_bytes")]


class DvgpuMemoryInfo(ctypes.Structure):
    _fields_ = [("free_bytes", ctypes.c_uint64), ("total_bytes", ctypes.c_uint64)]


class DvgpuResidentOp(ctypes.Structure):
    _fields_ = [("opcode", ctypes.c_uint32), ("kind", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("a", ctypes.c_uint64), ("b", ctypes.c_uint64),
                ("c", ctypes.c_uint64), ("a_offset", ctypes.c_uint64), ("b_offset", ctypes.c_uint64),
                ("c_offset", ctypes.c_uint64), ("length", ctypes.c_uint64), ("rows", ctypes.c_uint64),
                ("cols", ctypes.c_uint64), ("input", c_void_p), ("output", c_void_p)]


class DvgpuResidentStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "h2d_bytes", "d2h_bytes", "bytes_allocated", "fr_transform_elements",
        "g1_transform_elements", "msm_g1_points", "msm_g2_points", "calls")]


class DvgpuResidentDiagnostics(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "resident_allocations", "temporary_allocations", "temporary_peak_bytes",
        "journal_saved_bytes", "journal_saved_ranges", "journal_peak_bytes",
        "device_operation_ns", "device_operations", "physical_h2d_bytes", "physical_d2h_bytes",
        "explicit_d2d_bytes", "journal_metadata_capacity_bytes")]


class DvgpuResidentProfile(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "native_wall_ns", "journal_host_ns", "conversion_host_ns", "operation_host_ns",
        "wait_host_ns", "event_total_ns", "event_journal_ns", "event_operation_ns", "event_mask",
        "h2d_bytes", "d2h_bytes", "d2d_bytes")]


class DvgpuMsmTimings(ctypes.Structure):
    _fields_ = [("conversion_ns", ctypes.c_uint64), ("icicle_ns", ctypes.c_uint64),
                ("output_ns", ctypes.c_uint64)]


class DvgpuMsmTimeline(ctypes.Structure):
    _fields_ = [("wall_ns", ctypes.c_uint64 * 9), ("event_ns", ctypes.c_uint64 * 9),
                ("calls", ctypes.c_uint64 * 9), ("event_calls", ctypes.c_uint64 * 9),
                ("bytes", ctypes.c_uint64 * 9), ("kernel_ns", ctypes.c_uint64 * 10),
                ("kernel_calls", ctypes.c_uint64 * 10),
                ("instrumentation_inclusive_ns", ctypes.c_uint64),
                ("instrumentation_total_ns", ctypes.c_uint64), ("c", ctypes.c_uint64),
                ("bitsize", ctypes.c_uint64), ("windows", ctypes.c_uint64),
                ("buckets", ctypes.c_uint64)]


class DvgpuG2KernelConfig(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_uint32), ("schedule", ctypes.c_uint32),
                ("layout", ctypes.c_uint32), ("fused", ctypes.c_uint32)]


class DvgpuG2ExperimentStats(ctypes.Structure):
    _fields_ = [("repack_device_ns", ctypes.c_uint64), ("subgroup_kernel_ns", ctypes.c_uint64),
                ("logical_repack_bytes", ctypes.c_uint64)]


class DvgpuG2KernelResources(ctypes.Structure):
    _fields_ = [("registers", ctypes.c_uint32), ("block_threads", ctypes.c_uint32),
                ("active_blocks", ctypes.c_uint32), ("max_threads_per_sm", ctypes.c_uint32),
                ("local_bytes", ctypes.c_uint64), ("shared_bytes", ctypes.c_uint64),
                ("normalize_registers", ctypes.c_uint32), ("normalize_active_blocks", ctypes.c_uint32),
                ("normalize_local_bytes", ctypes.c_uint64), ("normalize_shared_bytes", ctypes.c_uint64)]


class DvgpuTuning(ctypes.Structure):
    _fields_ = [("msm_window_bits", ctypes.c_int32 * 25), ("ntt_pass_stages", ctypes.c_int32 * 21),
                ("msm_task_points", ctypes.c_int32), ("msm_chunk", ctypes.c_int32),
                ("msm_device_finalize_min", ctypes.c_int32), ("msm_chunk_by_log2", ctypes.c_int32 * 25)]


class DvgpuMsmItem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("point_base", ctypes.c_uint32), ("scalar_base", ctypes.c_uint32),
                ("window_bits", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class DvgpuNttRows(ctypes.Structure):
    _fields_ = [("in_", c_void_p), ("out", c_void_p), ("length", ctypes.c_uint64), ("rows", ctypes.c_uint64),
                ("inverse", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class DvgpuBatchSpec(ctypes.Structure):
    _fields_ = [("g1_points", c_void_p), ("g1_index", c_void_p), ("g1_scalars", c_void_p),
                ("g1_items", ctypes.POINTER(DvgpuMsmItem)), ("g1_count", ctypes.c_uint64),
                ("g2_points", c_void_p), ("g2_index", c_void_p), ("g2_scalars", c_void_p),
                ("g2_items", ctypes.POINTER(DvgpuMsmItem)), ("g2_count", ctypes.c_uint64),
                ("ntt", ctypes.POINTER(DvgpuNttRows)), ("ntt_count", ctypes.c_uint64),
                ("g2_wire", c_void_p), ("g2_wire_n", ctypes.c_uint64),
                ("g2_algorithm", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("g2_masks", c_void_p), ("root", c_void_p)]


class DvgpuUpdate(ctypes.Structure):
    _fields_ = [("srs_row", ctypes.c_uint32), ("opening_row", ctypes.c_uint32),
                ("commitment_row", ctypes.c_uint32), ("opening_cache_row", ctypes.c_uint32),
                ("coeff_offset", ctypes.c_uint64), ("eval_offset", ctypes.c_uint64),
                ("linear_offset", ctypes.c_uint64), ("value_slot", ctypes.c_uint64),
                ("log_n", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("indices", c_void_p), ("deltas", c_void_p)]


class DvgpuKzgTimings(ctypes.Structure):
    _fields_ = [("scalar_and_basis_ms", ctypes.c_float), ("convolution_ms", ctypes.c_float),
                ("domain_transform_ms", ctypes.c_float)]


class DvgpuPolyStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_uint64), ("products", ctypes.c_uint64), ("device_products", ctypes.c_uint64)]


_I = ctypes.c_int
_E = (ctypes.c_char_p, c_size_t)  # error, capacity

# name -> (restype, argtypes); every symbol the C headers declare.
SIGNATURES = {
    "dvgpu_backend_create": (_I, [ctypes.POINTER(c_void_p), *_E]),
    "dvgpu_backend_create_with_large_bucket_factor": (_I, [ctypes.c_uint32, ctypes.POINTER(c_void_p), *_E]),
    "dvgpu_backend_destroy": (None, [c_void_p]),
    "dvgpu_backend_device_identity": (_I, [c_void_p, ctypes.c_char_p, c_size_t, *_E]),
    "dvgpu_bn254_msm": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, *_E]),
    "dvgpu_bn254_g2_msm": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, *_E]),
    "dvgpu_bn254_ntt": (_I, [c_void_p, c_void_p, c_size_t, _I, c_void_p, c_void_p, *_E]),
    "dvgpu_bn254_msm_batch": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, *_E]),
    "dvgpu_setup_mode": (_I, [c_void_p, _I]),
    "dvgpu_setup_stats": (_I, [c_void_p, ctypes.POINTER(DvgpuSetupStats)]),
    "dvgpu_bn254_g2_validation_inspect": (_I, [c_void_p, c_void_p, c_size_t, ctypes.c_uint32, c_void_p,
                                               ctypes.POINTER(DvgpuG2ValidationReport), *_E]),
    "dvgpu_g2_validation_stats": (_I, [c_void_p, ctypes.POINTER(DvgpuG2ValidationStats)]),
    "dvgpu_g2_validation_traffic": (_I, [c_void_p, ctypes.POINTER(DvgpuG2ValidationTraffic)]),
    "dvgpu_bn254_fixed_base_g2_checked": (_I, [c_void_p, c_void_p, c_size_t, ctypes.c_uint32, c_void_p,
                                               ctypes.POINTER(DvgpuG2ValidationReport), *_E]),
    "dvgpu_backend_batch_stats": (_I, [c_void_p, ctypes.POINTER(DvgpuBatchStats)]),
    "dvgpu_backend_memory_info": (_I, [c_void_p, ctypes.POINTER(DvgpuMemoryInfo), *_E]),
    "dvgpu_resident_configure": (_I, [c_void_p, ctypes.c_uint64, ctypes.c_uint64, c_void_p, *_E]),
    "dvgpu_resident_call": (_I, [c_void_p, ctypes.POINTER(DvgpuResidentOp), c_u64p, *_E]),
    "dvgpu_resident_seal": (None, [c_void_p]),
    "dvgpu_resident_commit": (None, [c_void_p]),
    "dvgpu_resident_stats": (_I, [c_void_p, ctypes.POINTER(DvgpuResidentStats)]),
    "dvgpu_resident_diagnostics": (_I, [c_void_p, ctypes.POINTER(DvgpuResidentDiagnostics)]),
    "dvgpu_resident_call_profiled": (_I, [c_void_p, ctypes.POINTER(DvgpuResidentOp), c_u64p,
                                          ctypes.POINTER(DvgpuResidentProfile), *_E]),
    "dvgpu_bn254_msm_profiled": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p,
                                      ctypes.POINTER(DvgpuMsmTimings), *_E]),
    "dvgpu_bn254_msm_timeline": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, ctypes.c_char_p,
                                      ctypes.POINTER(DvgpuMsmTimings), ctypes.POINTER(DvgpuMsmTimeline), *_E]),
    "dvgpu_reserve_batch": (_I, [c_void_p, c_size_t, *_E]),
    "dvgpu_bn254_fixed_base": (_I, [c_void_p, c_void_p, c_size_t, _I, c_void_p, *_E]),
    "dvgpu_g2_psi_probe": (_I, [c_void_p, c_void_p, c_size_t, c_void_p, *_E]),
    "dvgpu_g2_test_corrupt_next": (_I, [c_void_p, ctypes.c_uint64]),
    "dvgpu_g2_kernel_config_set": (_I, [c_void_p, ctypes.POINTER(DvgpuG2KernelConfig)]),
    "dvgpu_g2_kernel_config_get": (_I, [c_void_p, ctypes.POINTER(DvgpuG2KernelConfig)]),
    "dvgpu_g2_kernel_resources": (_I, [c_void_p, ctypes.c_uint32, ctypes.POINTER(DvgpuG2KernelResources)]),
    "dvgpu_g2_experiment_stats": (_I, [c_void_p, ctypes.POINTER(DvgpuG2ExperimentStats)]),
    "dvgpu_bn254_tower_inspect": (_I, [c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_size_t,
                                       c_void_p, c_size_t, *_E]),
    "dvgpu_bn254_lines_inspect": (_I, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p,
                                       c_size_t, c_void_p, c_size_t, *_E]),
    "dvgpu_bn254_miller_inspect": (_I, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p,
                                        c_size_t, c_void_p, c_size_t, *_E]),
    # dynaverse_hip_ext.h
    "dvgpu_ext_alloc": (_I, [c_void_p, c_size_t, ctypes.POINTER(c_void_p), *_E]),
    "dvgpu_ext_free": (_I, [c_void_p, c_void_p]),
    "dvgpu_ext_copy_to_device": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, *_E]),
    "dvgpu_ext_copy_to_host": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, *_E]),
    "dvgpu_ext_stream": (_I, [c_void_p, ctypes.POINTER(c_void_p)]),
    "dvgpu_ext_device_copy": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, *_E]),
    "dvgpu_ext_set_tuning": (_I, [c_void_p, ctypes.POINTER(DvgpuTuning)]),
    "dvgpu_ext_get_tuning": (_I, [c_void_p, ctypes.POINTER(DvgpuTuning)]),
    "dvgpu_ext_msm_window": (_I, [c_void_p, ctypes.c_uint64]),
    "dvgpu_ext_msm_batch_device": (_I, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(DvgpuMsmItem),
                                        c_size_t, _I, c_void_p, ctypes.POINTER(ctypes.c_float), *_E]),
    "dvgpu_ext_msm_batch_finalize": (_I, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_size_t,
                                          ctypes.POINTER(DvgpuMsmItem), c_size_t, _I, c_void_p, c_void_p,
                                          ctypes.POINTER(ctypes.c_uint32), c_void_p, c_size_t, *_E]),
    "dvgpu_ext_msm_batch_finalize_noalloc": (_I, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p,
                                                  c_size_t, ctypes.POINTER(DvgpuMsmItem), c_size_t, _I, c_void_p,
                                                  c_void_p, *_E]),
    "dvgpu_ext_ntt_batch_device": (_I, [c_void_p, ctypes.POINTER(DvgpuNttRows), c_size_t, c_void_p,
                                        ctypes.POINTER(ctypes.c_float), *_E]),
    "dvgpu_ext_update_batch": (_I, [c_void_p, c_void_p, c_size_t, _I, c_void_p, c_size_t,
                                    ctypes.POINTER(DvgpuUpdate), c_size_t, c_void_p, c_void_p,
                                    ctypes.POINTER(ctypes.c_float), *_E]),
    "dvgpu_ext_update_reserve": (_I, [c_void_p, c_size_t, c_size_t, c_size_t, _I, c_void_p, *_E]),
    "dvgpu_ext_window_table_device": (_I, [c_void_p, c_void_p, c_size_t, _I, c_void_p, *_E]),
    # dynaverse_kzg.h
    "dvgpu_kzg_srs_create": (_I, [c_void_p, c_void_p, c_size_t, c_void_p, ctypes.POINTER(c_void_p), *_E]),
    "dvgpu_kzg_srs_destroy": (_I, [c_void_p, c_void_p]),
    "dvgpu_kzg_srs_basis": (_I, [c_void_p, c_void_p, c_u64p, c_u64p]),
    "dvgpu_kzg_precompute_single_openings": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_size_t,
                                                  c_void_p, c_void_p, ctypes.POINTER(DvgpuKzgTimings), *_E]),
    "dvgpu_ext_glv_split": (_I, [c_void_p, c_size_t, c_void_p]),
    "dvgpu_ext_glv_jsf": (_I, [c_void_p, c_size_t, c_void_p]),
    # dynaverse_pairing.h
    "dvgpu_bn254_pairing_check": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, c_void_p,
                                       ctypes.POINTER(ctypes.c_float), *_E]),
    "dvgpu_bn254_pairing_equations": (_I, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, ctypes.c_int,
                                           c_void_p, ctypes.POINTER(ctypes.c_float), *_E]),
    # dynaverse_poly.h
    "dvgpu_poly_mul": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_size_t,
                            ctypes.POINTER(c_size_t), *_E]),
    "dvgpu_poly_divide_exact": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p,
                                     c_size_t, ctypes.POINTER(c_size_t), *_E]),
    "dvgpu_poly_tree_create": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, ctypes.POINTER(c_void_p),
                                    ctypes.POINTER(DvgpuPolyStats), *_E]),
    "dvgpu_poly_tree_destroy": (_I, [c_void_p, c_void_p]),
    "dvgpu_poly_tree_vanishing": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, ctypes.POINTER(c_size_t), *_E]),
    "dvgpu_poly_tree_inverse_derivatives": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, ctypes.POINTER(c_size_t),
                                                 *_E]),
    "dvgpu_poly_tree_eval": (_I, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p,
                                  ctypes.POINTER(DvgpuPolyStats), *_E]),
    "dvgpu_poly_tree_interpolate": (_I, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t,
                                         ctypes.POINTER(c_size_t), ctypes.POINTER(DvgpuPolyStats), *_E]),
    "dvgpu_poly_tree_cauchy_sums": (_I, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p,
                                         ctypes.POINTER(DvgpuPolyStats), *_E]),
    "dvgpu_bn254_opening_cache_check": (_I, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p,
                                             c_size_t, c_void_p, c_void_p, c_void_p, ctypes.POINTER(ctypes.c_float),
                                             *_E]),
    "dvgpu_bn254_kzg_verify_batch": (_I, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_size_t, c_void_p, ctypes.POINTER(ctypes.c_float), *_E]),
    "dvgpu_ext_batch_create": (_I, [c_void_p, ctypes.POINTER(DvgpuBatchSpec), _I, ctypes.POINTER(c_void_p), *_E]),
    "dvgpu_ext_batch_run": (_I, [c_void_p, c_void_p, c_void_p, c_void_p,
                                 ctypes.POINTER(DvgpuG2ValidationReport), ctypes.POINTER(ctypes.c_float), *_E]),
    "dvgpu_ext_batch_destroy": (_I, [c_void_p, c_void_p]),
    "dvgpu_ext_fixed_base_device": (_I, [c_void_p, c_void_p, c_size_t, _I, c_void_p, *_E]),
    "dvgpu_ext_points_to_mont": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, _I, *_E]),
    "dvgpu_ext_msm_device": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, _I, _I, c_void_p,
                                  ctypes.POINTER(ctypes.c_float), *_E]),
    "dvgpu_ext_ntt_device": (_I, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, _I, c_void_p,
                                  ctypes.POINTER(ctypes.c_float), *_E]),
}

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libdynaverse_hip.so is not built (run __graft_entry__.build()); "
                "this package has no CPU fallback")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
