"""HipBackend: the operator backend over the MI355X library.

Mirrors the reference's ``IcicleBackend`` (crates/gpu/src/icicle.rs:98-457)
method for method -- ``msm``, ``msm_batch``, ``msm_sparse_batch``,
``msm_g2_host``, ``ntt``, ``fixed_base_g1/g2`` -- with the same argument
meaning, validation order and error behaviour: inputs are checked before any
native call, a native failure raises ``BackendError`` carrying the status, and
an execution failure retires the native owner (later calls raise status -4).
There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .codec import (FR_MODULUS, G1_BYTES, G2_BYTES, MAX_MSM_BATCH_SIZE, MAX_MSM_SIZE,
                    MAX_NTT_SIZE, decode_fr, decode_g1, decode_g2, encode_fr, encode_g1,
                    encode_g2, g1_on_curve, g2_on_curve, root_of_unity_2_20)


class BackendError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"HIP backend error {status}: {message}")
        self.status = status
        self.message = message


@dataclass
class BatchMsmRequest:
    """crates/protocol/src/types.rs:253-259: point sets sharing scalars."""
    point_sets: Sequence[Sequence]
    shared_scalars: Sequence[int]
    offsets: Sequence[int]
    lengths: Sequence[int]


@dataclass
class SparseMsmRequest:
    """crates/protocol/src/types.rs:261-277: sum value * bases[index]."""
    bases: Sequence
    entries: Sequence  # (index, value) pairs; order/repeats/zeros preserved

    def validate(self):
        if not all(0 <= i < len(self.bases) for i, _ in self.entries):
            raise ValueError("sparse MSM index out of range")


def _err():
    return ctypes.create_string_buffer(256)


def _check(status: int, err) -> None:
    if status != 0:
        raise BackendError(status, err.value.decode(errors="replace"))


def _ptr(buf) -> ctypes.c_void_p:
    if isinstance(buf, np.ndarray):
        return ctypes.c_void_p(buf.ctypes.data)
    return ctypes.cast(ctypes.c_char_p(buf), ctypes.c_void_p)


class DeviceArray:
    """Device memory owned by a HipBackend owner (freed with it)."""

    def __init__(self, backend: "HipBackend", nbytes: int):
        self.backend = backend
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        err = _err()
        _check(backend.lib.dvgpu_ext_alloc(backend.handle, self.nbytes, ctypes.byref(p), err, 256), err)
        self.ptr = p.value

    def upload(self, arr) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("upload larger than allocation")
        err = _err()
        _check(self.backend.lib.dvgpu_ext_copy_to_device(self.backend.handle, self.ptr,
                                                         ctypes.c_void_p(arr.ctypes.data), arr.nbytes,
                                                         err, 256), err)

    def download(self, nbytes: Optional[int] = None) -> np.ndarray:
        n = self.nbytes if nbytes is None else int(nbytes)
        out = np.empty(n, dtype=np.uint8)
        err = _err()
        _check(self.backend.lib.dvgpu_ext_copy_to_host(self.backend.handle, ctypes.c_void_p(out.ctypes.data),
                                                       self.ptr, n, err, 256), err)
        return out

    def upload_at(self, offset: int, arr) -> None:
        """Copy host bytes to [offset, offset + len) of this allocation."""
        arr = np.ascontiguousarray(np.frombuffer(bytes(arr), dtype=np.uint8) if not isinstance(arr, np.ndarray)
                                   else arr)
        if offset < 0 or offset + arr.nbytes > self.nbytes:
            raise ValueError("upload outside the allocation")
        if not arr.nbytes:
            return
        err = _err()
        _check(self.backend.lib.dvgpu_ext_copy_to_device(self.backend.handle, self.ptr + offset,
                                                         ctypes.c_void_p(arr.ctypes.data), arr.nbytes,
                                                         err, 256), err)

    def download_at(self, offset: int, nbytes: int) -> np.ndarray:
        if offset < 0 or offset + nbytes > self.nbytes:
            raise ValueError("download outside the allocation")
        out = np.empty(nbytes, dtype=np.uint8)
        err = _err()
        _check(self.backend.lib.dvgpu_ext_copy_to_host(self.backend.handle, ctypes.c_void_p(out.ctypes.data),
                                                       self.ptr + offset, nbytes, err, 256), err)
        return out

    def free(self) -> None:
        if self.ptr and self.backend.handle is not None and self.backend.handle.value:
            self.backend.lib.dvgpu_ext_free(self.backend.handle, self.ptr)
        self.ptr = None


class HipBackend:
    """One native owner on HIP device 0 (not shareable across threads)."""

    def __init__(self, large_bucket_factor: int = 0):
        if not 0 <= large_bucket_factor <= 100:
            raise ValueError("large bucket factor must be within 0..=100")
        self.lib = _lib.load()
        handle = ctypes.c_void_p()
        err = _err()
        _check(self.lib.dvgpu_backend_create_with_large_bucket_factor(
            large_bucket_factor, ctypes.byref(handle), err, 256), err)
        if not handle.value:
            raise BackendError(-2, "native backend returned a null owner")
        self.handle = handle
        self.large_bucket_factor = large_bucket_factor

    # ---- lifecycle ------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.dvgpu_backend_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def device_identity(self) -> str:
        out = ctypes.create_string_buffer(320)
        err = _err()
        _check(self.lib.dvgpu_backend_device_identity(self.handle, out, 320, err, 256), err)
        return out.value.decode()

    def memory_info(self):
        info = _lib.DvgpuMemoryInfo()
        err = _err()
        _check(self.lib.dvgpu_backend_memory_info(self.handle, ctypes.byref(info), err, 256), err)
        return info.free_bytes, info.total_bytes

    def batch_stats(self):
        st = _lib.DvgpuBatchStats()
        if self.lib.dvgpu_backend_batch_stats(self.handle, ctypes.byref(st)) != 0:
            raise BackendError(-1, "invalid native batch stats owner")
        return {k: getattr(st, k) for k, _ in st._fields_}

    # ---- raw encoded calls (caller-validated bytes, as the C ABI) ---------
    def msm_encoded(self, points: bytes, scalars: bytes, n: int) -> bytes:
        out = ctypes.create_string_buffer(64)
        err = _err()
        _check(self.lib.dvgpu_bn254_msm(self.handle, _ptr(points), _ptr(scalars), n, out, err, 256), err)
        return out.raw

    def g2_msm_encoded(self, points: bytes, scalars: bytes, n: int) -> bytes:
        out = ctypes.create_string_buffer(128)
        err = _err()
        _check(self.lib.dvgpu_bn254_g2_msm(self.handle, _ptr(points), _ptr(scalars), n, out, err, 256), err)
        return out.raw

    def msm_batch_encoded(self, points: bytes, scalars: bytes, length: int, count: int) -> bytes:
        out = ctypes.create_string_buffer(64 * count)
        err = _err()
        _check(self.lib.dvgpu_bn254_msm_batch(self.handle, _ptr(points), _ptr(scalars), length, count,
                                              out, err, 256), err)
        return out.raw

    def ntt_encoded(self, values: bytes, n: int, inverse: bool, root: Optional[bytes] = None) -> bytes:
        root = root if root is not None else encode_fr(root_of_unity_2_20())
        out = ctypes.create_string_buffer(32 * n)
        err = _err()
        _check(self.lib.dvgpu_bn254_ntt(self.handle, _ptr(values), n, int(inverse), _ptr(root), out,
                                        err, 256), err)
        return out.raw

    # ---- OperatorBackend (types.rs:281-359) -------------------------------
    def msm(self, points: Sequence, scalars: Sequence[int]):
        """icicle.rs:209-274: length/size checks, on-curve check, encode."""
        if len(points) != len(scalars):
            raise ValueError("MSM length mismatch")
        if len(points) > MAX_MSM_SIZE:
            raise ValueError("MSM exceeds the ABI bound of 2^20 points")
        if not points:
            return None
        for p in points:
            if not g1_on_curve(p):
                raise ValueError("invalid G1 point")
        pts = b"".join(encode_g1(p) for p in points)
        scs = b"".join(encode_fr(s) for s in scalars)
        return decode_g1(self.msm_encoded(pts, scs, len(points)))

    def msm_profiled(self, points, scalars):
        if len(points) != len(scalars):
            raise ValueError("MSM length mismatch")
        if not points:
            return None, {"conversion_ns": 0, "icicle_ns": 0, "output_ns": 0}
        for p in points:
            if not g1_on_curve(p):
                raise ValueError("invalid G1 point")
        pts = b"".join(encode_g1(p) for p in points)
        scs = b"".join(encode_fr(s) for s in scalars)
        out = ctypes.create_string_buffer(64)
        t = _lib.DvgpuMsmTimings()
        err = _err()
        _check(self.lib.dvgpu_bn254_msm_profiled(self.handle, _ptr(pts), _ptr(scs), len(points), out,
                                                 ctypes.byref(t), err, 256), err)
        return decode_g1(out.raw), {k: getattr(t, k) for k, _ in t._fields_}

    def msm_timeline(self, points, scalars, label: str = "msm"):
        if len(points) != len(scalars):
            raise ValueError("MSM length mismatch")
        for p in points:
            if not g1_on_curve(p):
                raise ValueError("invalid G1 point")
        pts = b"".join(encode_g1(p) for p in points) or b"\0"
        scs = b"".join(encode_fr(s) for s in scalars) or b"\0"
        out = ctypes.create_string_buffer(64)
        t = _lib.DvgpuMsmTimings()
        tl = _lib.DvgpuMsmTimeline()
        err = _err()
        _check(self.lib.dvgpu_bn254_msm_timeline(self.handle, _ptr(pts), _ptr(scs), len(points), out,
                                                 label.encode(), ctypes.byref(t), ctypes.byref(tl),
                                                 err, 256), err)
        return decode_g1(out.raw), {k: getattr(t, k) for k, _ in t._fields_}, tl

    def _submit_encoded_batches(self, lengths: List[int], encode) -> list:
        """icicle.rs:281-339: bin by next power of two, <= 64 sets and
        <= 2^20 padded points per native call; empty sets are identity."""
        results = [None] * len(lengths)
        bins = {}
        for i, ln in enumerate(lengths):
            if ln > MAX_MSM_SIZE:
                raise ValueError("MSM exceeds the ABI bound of 2^20 points")
            if ln > 0:
                bins.setdefault((ln - 1).bit_length(), []).append(i)
        for b in sorted(bins):
            padded = 1 << b
            chunk = min(MAX_MSM_BATCH_SIZE, MAX_MSM_SIZE // padded)
            idxs = bins[b]
            for g in range(0, len(idxs), chunk):
                group = idxs[g:g + chunk]
                pts = bytearray(padded * len(group) * G1_BYTES)
                scs = bytearray(padded * len(group) * 32)
                for slot, index in enumerate(group):
                    p_enc, s_enc = encode(index)
                    pts[slot * padded * 64: slot * padded * 64 + len(p_enc)] = p_enc
                    scs[slot * padded * 32: slot * padded * 32 + len(s_enc)] = s_enc
                out = self.msm_batch_encoded(bytes(pts), bytes(scs), padded, len(group))
                for slot, index in enumerate(group):
                    results[index] = decode_g1(out[slot * 64:(slot + 1) * 64])
        return results

    def msm_batch(self, request: BatchMsmRequest) -> list:
        count = len(request.point_sets)
        if len(request.offsets) != count or len(request.lengths) != count:
            raise ValueError("batch shape mismatch")
        for i in range(count):
            ln = request.lengths[i]
            end = request.offsets[i] + ln
            if end > len(request.shared_scalars):
                raise ValueError("scalar range out of bounds")
            if ln > len(request.point_sets[i]):
                raise ValueError("point range out of bounds")
            if ln > MAX_MSM_SIZE:
                raise ValueError("MSM exceeds the ABI bound of 2^20 points")
            for p in request.point_sets[i][:ln]:
                if not g1_on_curve(p):
                    raise ValueError("invalid G1 point")

        def encode(i):
            ln, off = request.lengths[i], request.offsets[i]
            return (b"".join(encode_g1(p) for p in request.point_sets[i][:ln]),
                    b"".join(encode_fr(s) for s in request.shared_scalars[off:off + ln]))
        return self._submit_encoded_batches(list(request.lengths), encode)

    def msm_sparse_batch(self, requests: Sequence[SparseMsmRequest]) -> list:
        """icicle.rs:375-397: selected bases encoded directly, no gather."""
        lengths = []
        for r in requests:
            if len(r.entries) > MAX_MSM_SIZE:
                raise ValueError("MSM exceeds the ABI bound of 2^20 points")
            r.validate()
            for idx, _ in r.entries:
                if not g1_on_curve(r.bases[idx]):
                    raise ValueError("invalid G1 point")
            lengths.append(len(r.entries))

        def encode(i):
            r = requests[i]
            return (b"".join(encode_g1(r.bases[idx]) for idx, _ in r.entries),
                    b"".join(encode_fr(v) for _, v in r.entries))
        return self._submit_encoded_batches(lengths, encode)

    def msm_g2_host(self, points: Sequence, scalars: Sequence[int]):
        """icicle.rs:403-425 (on-curve checked here; subgroup membership is
        checked on the device by ``inspect_g2_validation`` when requested)."""
        if len(points) != len(scalars):
            raise ValueError("G2 MSM length mismatch")
        if len(points) > MAX_MSM_SIZE:
            raise ValueError("G2 MSM exceeds the ABI bound of 2^20 points")
        for p in points:
            if not g2_on_curve(p):
                raise ValueError("invalid G2 point")
        if not points:
            return None
        pts = b"".join(encode_g2(p) for p in points)
        scs = b"".join(encode_fr(s) for s in scalars)
        return decode_g2(self.g2_msm_encoded(pts, scs, len(points)))

    def ntt(self, values: List[int], inverse: bool) -> None:
        """icicle.rs:429-456: in place; caller's values untouched on error."""
        n = len(values)
        if n == 0 or n & (n - 1):
            raise ValueError("NTT size is not a power of two")
        if n > MAX_NTT_SIZE:
            raise ValueError("NTT exceeds the ABI bound of 2^20 points")
        enc = b"".join(encode_fr(v) for v in values)
        out = self.ntt_encoded(enc, n, inverse)
        decoded = [decode_fr(out[32 * i:32 * i + 32]) for i in range(n)]
        values[:] = decoded

    def gpu_capabilities(self):
        return {"g1_msm": True, "fr_ntt": True, "g2_msm": True}

    # ---- device-resident extension (dynaverse_hip_ext.h) -------------------
    def alloc(self, nbytes: int) -> "DeviceArray":
        return DeviceArray(self, nbytes)

    def to_device(self, data) -> "DeviceArray":
        arr = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8)
                                   if not isinstance(data, np.ndarray) else data)
        d = DeviceArray(self, arr.nbytes)
        d.upload(arr)
        return d

    def device_copy(self, dst_ptr: int, src_ptr: int, nbytes: int) -> None:
        err = _err()
        _check(self.lib.dvgpu_ext_device_copy(self.handle, dst_ptr, src_ptr, nbytes, err, 256), err)

    def stream_handle(self) -> int:
        s = ctypes.c_void_p()
        _check(self.lib.dvgpu_ext_stream(self.handle, ctypes.byref(s)), _err())
        return s.value or 0

    def points_to_mont_device(self, canonical_ptr: int, mont_ptr: int, n: int, g2: bool = False):
        err = _err()
        _check(self.lib.dvgpu_ext_points_to_mont(self.handle, canonical_ptr, mont_ptr, n, int(g2), err, 256), err)

    def msm_device(self, points_ptr: int, scalars_ptr: int, n: int, g2: bool = False,
                   window_bits: int = 0, phases: bool = False):
        out = ctypes.create_string_buffer(128 if g2 else 64)
        ph = (ctypes.c_float * 7)() if phases else None
        err = _err()
        _check(self.lib.dvgpu_ext_msm_device(self.handle, points_ptr, scalars_ptr, n, int(g2), window_bits,
                                             out, ph, err, 256), err)
        res = out.raw
        return (res, list(ph)) if phases else res

    def ntt_device(self, in_ptr: int, out_ptr: int, length: int, rows: int, inverse: bool,
                   timed: bool = False):
        root = encode_fr(root_of_unity_2_20())
        ms = ctypes.c_float(0.0)
        err = _err()
        _check(self.lib.dvgpu_ext_ntt_device(self.handle, in_ptr, out_ptr, length, rows, int(inverse),
                                             _ptr(root), ctypes.byref(ms) if timed else None, err, 256), err)
        return ms.value if timed else None


# ---- fixed-base setup and G2 validation (setup.rs, g2_validation.rs) ------

SETUP_CHUNK = 4096
G2_ALGORITHMS = {"full_order": 1, "frobenius127": 2, "frobenius63": 3}


def _setup_methods():
    def set_setup_mode(self, mode: int) -> None:
        """setup.rs:82-86: 0 ordinary, 1 table, 2 table + batch normalize."""
        if self.lib.dvgpu_setup_mode(self.handle, int(mode)) != 0:
            raise BackendError(-1, "invalid or retired setup owner")

    def setup_stats(self):
        st = _lib.DvgpuSetupStats()
        if self.lib.dvgpu_setup_stats(self.handle, ctypes.byref(st)) != 0:
            raise BackendError(-1, "invalid setup stats owner")
        return {"table_builds": list(st.table_builds), "calls": list(st.calls),
                "points": list(st.points), "table_bytes": st.table_bytes,
                "device_ns": list(st.device_ns), "native_wall_ns": st.native_wall_ns}

    def fixed_base_encoded(self, scalars: bytes, n: int, g2: bool) -> bytes:
        out = ctypes.create_string_buffer(n * (128 if g2 else 64))
        err = _err()
        _check(self.lib.dvgpu_bn254_fixed_base(self.handle, _ptr(scalars), n, int(g2), out, err, 256), err)
        return out.raw

    def fixed_base_g1(self, scalars: Sequence[int]) -> list:
        """setup.rs:113-126: 4096-scalar chunks, decoded with the on-curve check."""
        out = []
        for c in range(0, len(scalars), SETUP_CHUNK):
            chunk = scalars[c:c + SETUP_CHUNK]
            raw = self.fixed_base_encoded(b"".join(encode_fr(s) for s in chunk), len(chunk), False)
            out.extend(decode_g1(raw[64 * i:64 * i + 64]) for i in range(len(chunk)))
        return out

    def fixed_base_g2(self, scalars: Sequence[int], algorithm: int = 3) -> list:
        """setup.rs:128-143 with device validation (g2_validation.rs:149-169):
        every chunk is validated on the GPU before it is published."""
        out = []
        for c in range(0, len(scalars), SETUP_CHUNK):
            chunk = scalars[c:c + SETUP_CHUNK]
            raw, report = self.fixed_base_g2_checked_encoded(
                b"".join(encode_fr(s) for s in chunk), len(chunk), algorithm)
            if report["invalid_points"]:
                raise BackendError(-5, "G2 validation rejected index %d (mask %d)"
                                   % (c + report["first_invalid"], report["failure_mask"]))
            out.extend(decode_g2(raw[128 * i:128 * i + 128]) for i in range(len(chunk)))
        return out

    def fixed_base_g2_checked_encoded(self, scalars: bytes, n: int, algorithm: int):
        out = ctypes.create_string_buffer(max(1, n * 128))
        rep = _lib.DvgpuG2ValidationReport()
        err = _err()
        status = self.lib.dvgpu_bn254_fixed_base_g2_checked(self.handle, _ptr(scalars), n, algorithm, out,
                                                            ctypes.byref(rep), err, 256)
        report = {k: getattr(rep, k) for k, _ in rep._fields_}
        if status == -5:
            return None, report
        _check(status, err)
        return out.raw[:n * 128], report

    def inspect_g2_validation(self, wire: bytes, algorithm: int):
        """g2_validation.rs:191-206: masks (1 encoding, 2 curve, 4 subgroup)."""
        if len(wire) % 128 or len(wire) // 128 > SETUP_CHUNK or algorithm not in (1, 2, 3):
            raise ValueError("invalid G2 validation input")
        n = len(wire) // 128
        masks = (ctypes.c_uint32 * max(1, n))()
        rep = _lib.DvgpuG2ValidationReport()
        err = _err()
        _check(self.lib.dvgpu_bn254_g2_validation_inspect(self.handle, _ptr(wire or b"\0"), n, algorithm,
                                                          masks, ctypes.byref(rep), err, 256), err)
        return list(masks)[:n], {k: getattr(rep, k) for k, _ in rep._fields_}

    def inspect_g2_validation_batched(self, wire: bytes, algorithm: int):
        """g2_validation.rs:114-125: 4096-point blocks, global first_invalid."""
        masks = []
        comb = {"checked_points": 0, "invalid_points": 0, "first_invalid": (1 << 64) - 1,
                "failure_mask": 0, "algorithm": algorithm}
        for block, c in enumerate(range(0, len(wire), SETUP_CHUNK * 128)):
            m, r = self.inspect_g2_validation(wire[c:c + SETUP_CHUNK * 128], algorithm)
            comb["checked_points"] += r["checked_points"]
            comb["invalid_points"] += r["invalid_points"]
            comb["failure_mask"] |= r["failure_mask"]
            if r["invalid_points"]:
                comb["first_invalid"] = min(comb["first_invalid"], block * SETUP_CHUNK + r["first_invalid"])
            masks.extend(m)
        return masks, comb

    def _struct_getter(fn_name, cls):
        def get(self):
            st = cls()
            if getattr(self.lib, fn_name)(self.handle, ctypes.byref(st)) != 0:
                raise BackendError(-1, fn_name + " failed")
            return {k: getattr(st, k) for k, _ in st._fields_}
        return get

    def set_g2_kernel_config(self, threads=128, schedule=0, layout=0, fused=0):
        cfg = _lib.DvgpuG2KernelConfig(threads, schedule, layout, fused)
        status = self.lib.dvgpu_g2_kernel_config_set(self.handle, ctypes.byref(cfg))
        if status != 0:
            raise BackendError(status, "invalid or retired G2 kernel config")

    def g2_kernel_resources(self, algorithm: int):
        r = _lib.DvgpuG2KernelResources()
        status = self.lib.dvgpu_g2_kernel_resources(self.handle, algorithm, ctypes.byref(r))
        if status != 0:
            raise BackendError(status, "G2 kernel resource query failed")
        return {k: getattr(r, k) for k, _ in r._fields_}

    def g2_test_corrupt_next(self, index: int) -> int:
        return self.lib.dvgpu_g2_test_corrupt_next(self.handle, index)

    def g2_psi_probe(self, points: bytes, n: int) -> bytes:
        out = ctypes.create_string_buffer(max(1, n * 576))
        err = _err()
        _check(self.lib.dvgpu_g2_psi_probe(self.handle, _ptr(points or b"\0"), n, out, err, 256), err)
        return out.raw[:n * 576]

    def fixed_base_device(self, scalars_ptr: int, n: int, out_ptr: int, g2: bool = False):
        """[s_i] G from device scalars straight into device Montgomery points."""
        err = _err()
        _check(self.lib.dvgpu_ext_fixed_base_device(self.handle, scalars_ptr, n, int(g2), out_ptr, err, 256), err)

    ns = dict(set_setup_mode=set_setup_mode, setup_stats=setup_stats,
              fixed_base_encoded=fixed_base_encoded, fixed_base_g1=fixed_base_g1,
              fixed_base_g2=fixed_base_g2, fixed_base_g2_checked_encoded=fixed_base_g2_checked_encoded,
              inspect_g2_validation=inspect_g2_validation,
              inspect_g2_validation_batched=inspect_g2_validation_batched,
              g2_validation_stats=_struct_getter("dvgpu_g2_validation_stats", _lib.DvgpuG2ValidationStats),
              g2_validation_traffic=_struct_getter("dvgpu_g2_validation_traffic", _lib.DvgpuG2ValidationTraffic),
              g2_experiment_stats=_struct_getter("dvgpu_g2_experiment_stats", _lib.DvgpuG2ExperimentStats),
              set_g2_kernel_config=set_g2_kernel_config, g2_kernel_resources=g2_kernel_resources,
              g2_test_corrupt_next=g2_test_corrupt_next, g2_psi_probe=g2_psi_probe,
              fixed_base_device=fixed_base_device)
    for k, v in ns.items():
        setattr(HipBackend, k, v)


_setup_methods()
