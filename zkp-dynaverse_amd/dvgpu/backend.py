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
            raise ValueError("MSM exceeds baseline size bound (2^20)")
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
                raise ValueError("MSM exceeds baseline size bound (2^20)")
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
                raise ValueError("MSM exceeds baseline size bound (2^20)")
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
                raise ValueError("MSM exceeds baseline size bound (2^20)")
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
            raise ValueError("G2 MSM exceeds baseline size bound (2^20)")
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
            raise ValueError("NTT exceeds baseline size bound (2^20)")
        enc = b"".join(encode_fr(v) for v in values)
        out = self.ntt_encoded(enc, n, inverse)
        decoded = [decode_fr(out[32 * i:32 * i + 32]) for i in range(n)]
        values[:] = decoded

    def gpu_capabilities(self):
        return {"g1_msm": True, "fr_ntt": True, "g2_msm": True}

    # ---- device-resident extension (dynaverse_hip_ext.h) -------------------
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
