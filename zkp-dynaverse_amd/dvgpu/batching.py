"""Batching form: many uneven MSMs (G1, G2), NTT row batches and G2
validation as one launch sequence, optionally one hipGraph replay.

Window widths come from the tuning table (``autotune``) unless given per
item.  All inputs are device-resident (DeviceArray); results return to the
host in the canonical encoding.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

from . import _lib
from .backend import BackendError, DeviceArray, HipBackend, _check, _err
from .codec import encode_fr, root_of_unity_2_20


@dataclass
class MsmItem:
    n: int
    point_base: int = 0
    scalar_base: int = 0
    window_bits: int = 0


def _addr(x) -> int:
    return x.ptr if isinstance(x, DeviceArray) else int(x)


def _items(items: Sequence[MsmItem]):
    arr = (_lib.DvgpuMsmItem * max(1, len(items)))()
    for i, it in enumerate(items):
        arr[i] = _lib.DvgpuMsmItem(it.n, it.point_base, it.scalar_base, it.window_bits, 0)
    return arr


def msm_batch_device(backend: HipBackend, points: DeviceArray, scalars: DeviceArray, items: Sequence[MsmItem],
                     g2: bool = False, index: Optional[DeviceArray] = None, phases: bool = False):
    """One kernel sequence for all MSMs; returns canonical results per item."""
    arr = _items(items)
    width = 128 if g2 else 64
    out = ctypes.create_string_buffer(max(1, width * len(items)))
    ph = (ctypes.c_float * 7)() if phases else None
    err = _err()
    _check(backend.lib.dvgpu_ext_msm_batch_device(
        backend.handle, points.ptr, index.ptr if index else None, scalars.ptr, arr, len(items), int(g2),
        out, ph, err, 256), err)
    res = [out.raw[width * i:width * (i + 1)] for i in range(len(items))]
    return (res, list(ph)) if phases else res


def msm_batch_finalize(backend: HipBackend, points: DeviceArray, scalars: DeviceArray, items: Sequence[MsmItem],
                       g2: bool = False, index: Optional[DeviceArray] = None,
                       results: Optional[DeviceArray] = None, slots: Optional[Sequence[int]] = None,
                       slot_points: Optional[DeviceArray] = None, host: bool = True):
    """Window sums finished on the device.  `slots[i]` (or NO_SLOT) names the
    point of `slot_points` that MSM i's result replaces.  Returns the
    canonical results (host) when `host`."""
    arr = _items(items)
    width = 128 if g2 else 64
    aw = width
    out = ctypes.create_string_buffer(max(1, width * len(items))) if host else None
    sl = (ctypes.c_uint32 * len(items))(*slots) if slots is not None else None
    err = _err()
    _check(backend.lib.dvgpu_ext_msm_batch_finalize(
        backend.handle, points.ptr, points.nbytes // aw, index.ptr if index else None,
        index.nbytes // 4 if index else 0, scalars.ptr, scalars.nbytes // 32, arr, len(items), int(g2),
        results.ptr if results else None, out, sl, slot_points.ptr if slot_points else None,
        slot_points.nbytes // aw if slot_points else 0, err, 256), err)
    return [out.raw[width * i:width * (i + 1)] for i in range(len(items))] if host else None


def msm_batch_finalize_noalloc(backend: HipBackend, points: DeviceArray, scalars: DeviceArray,
                               items: Sequence[MsmItem], results: DeviceArray, g2: bool = False,
                               index: Optional[DeviceArray] = None, host: bool = True):
    """msm_batch_finalize without any device allocation: intermediates live in
    the library's fixed scratch array (NOALLOC_SCRATCH_BYTES per device), so
    batches that do not fit fail with BackendError (status -1) and calls
    overlapping on one device race.  `results` (device) is required."""
    arr = _items(items)
    width = 128 if g2 else 64
    out = ctypes.create_string_buffer(max(1, width * len(items))) if host else None
    err = _err()
    _check(backend.lib.dvgpu_ext_msm_batch_finalize_noalloc(
        backend.handle, points.ptr, points.nbytes // width, index.ptr if index else None,
        index.nbytes // 4 if index else 0, scalars.ptr, scalars.nbytes // 32, arr, len(items), int(g2),
        results.ptr, out, err, 256), err)
    return [out.raw[width * i:width * (i + 1)] for i in range(len(items))] if host else None


NOALLOC_SCRATCH_BYTES = 64 << 20
NO_SLOT = 0xFFFFFFFF


@dataclass
class NttRows:
    inp: object          # DeviceArray or a device address
    out: object
    length: int
    rows: int
    inverse: bool = False


def ntt_batch_device(backend: HipBackend, rows: Sequence[NttRows], timed: bool = False):
    """All row batches in one fused launch sequence (entries must not overlap)."""
    arr = (_lib.DvgpuNttRows * len(rows))()
    for i, r in enumerate(rows):
        arr[i] = _lib.DvgpuNttRows(_addr(r.inp), _addr(r.out), r.length, r.rows, int(r.inverse), 0)
    root = ctypes.create_string_buffer(encode_fr(root_of_unity_2_20()))
    ms = ctypes.c_float(0)
    err = _err()
    _check(backend.lib.dvgpu_ext_ntt_batch_device(backend.handle, arr, len(rows), root,
                                                  ctypes.byref(ms) if timed else None, err, 256), err)
    return ms.value if timed else None


class DeviceBatch:
    """A planned batch: G1 MSMs, G2 MSMs, NTT row batches, G2 validation."""

    def __init__(self, backend: HipBackend, *, g1=None, g2=None, ntt: Sequence[NttRows] = (),
                 g2_wire: Optional[DeviceArray] = None, g2_wire_n: int = 0, g2_algorithm: int = 3,
                 g2_masks: Optional[DeviceArray] = None, use_graph: bool = True):
        """g1 / g2: (points, scalars, items[, index]) with DeviceArrays."""
        self.b = backend
        self._keep = []
        spec = _lib.DvgpuBatchSpec()
        self.n1 = self.n2 = 0
        if g1:
            pts, sc, items = g1[:3]
            idx = g1[3] if len(g1) > 3 else None
            arr = _items(items)
            self._keep.append(arr)
            spec.g1_points, spec.g1_scalars = pts.ptr, sc.ptr
            spec.g1_index = idx.ptr if idx else None
            spec.g1_items, spec.g1_count = arr, len(items)
            self.n1 = len(items)
        if g2:
            pts, sc, items = g2[:3]
            idx = g2[3] if len(g2) > 3 else None
            arr = _items(items)
            self._keep.append(arr)
            spec.g2_points, spec.g2_scalars = pts.ptr, sc.ptr
            spec.g2_index = idx.ptr if idx else None
            spec.g2_items, spec.g2_count = arr, len(items)
            self.n2 = len(items)
        if ntt:
            rows = (_lib.DvgpuNttRows * len(ntt))()
            for i, r in enumerate(ntt):
                rows[i] = _lib.DvgpuNttRows(_addr(r.inp), _addr(r.out), r.length, r.rows, int(r.inverse), 0)
            self._keep.append(rows)
            spec.ntt, spec.ntt_count = rows, len(ntt)
        root = ctypes.create_string_buffer(encode_fr(root_of_unity_2_20()))
        self._keep.append(root)
        spec.root = ctypes.addressof(root)
        if g2_wire_n:
            spec.g2_wire, spec.g2_wire_n, spec.g2_algorithm = g2_wire.ptr, g2_wire_n, g2_algorithm
            spec.g2_masks = g2_masks.ptr
        self.wire_n = g2_wire_n
        h = ctypes.c_void_p()
        err = _err()
        _check(backend.lib.dvgpu_ext_batch_create(backend.handle, ctypes.byref(spec), int(use_graph),
                                                  ctypes.byref(h), err, 256), err)
        self.h = h

    def run(self):
        """-> (g1 results, g2 results, report dict or None, device ms)."""
        o1 = ctypes.create_string_buffer(max(1, 64 * self.n1))
        o2 = ctypes.create_string_buffer(max(1, 128 * self.n2))
        rep = _lib.DvgpuG2ValidationReport()
        ms = ctypes.c_float(0)
        err = _err()
        _check(self.b.lib.dvgpu_ext_batch_run(self.b.handle, self.h, o1, o2, ctypes.byref(rep), ctypes.byref(ms),
                                              err, 256), err)
        r1 = [o1.raw[64 * i:64 * i + 64] for i in range(self.n1)]
        r2 = [o2.raw[128 * i:128 * i + 128] for i in range(self.n2)]
        report = {k: getattr(rep, k) for k, _ in rep._fields_} if self.wire_n else None
        return r1, r2, report, ms.value

    def close(self):
        if self.h is not None and self.b.handle is not None and self.b.handle.value:
            self.b.lib.dvgpu_ext_batch_destroy(self.b.handle, self.h)
        self.h = None
