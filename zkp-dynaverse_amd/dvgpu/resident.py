"""Resident operator engine: Python mirror of the reference's RebuildBackend
over the resident C ABI (crates/gpu/src/resident.rs:249-418,
crates/protocol/src/incremental_rebuild/operators.rs:52-111).

Buffers are owner-tagged handles; Fr values cross as canonical ints, G1
points as affine (x, y) / None.  Transactions journal each element's first
write; rollback restores it.
"""
import ctypes
from typing import List, Optional, Sequence

from . import _lib
from .backend import BackendError, HipBackend, _check, _err
from .codec import decode_g1, decode_g2, encode_fr, encode_g1, encode_g2, root_of_unity_2_20

ALLOC, LENGTH, KIND, BEGIN, ROLLBACK, WRITE, READ, COPY, REVERSE, TRANSPOSE, TRANSFORM, ZIP, SCALE, \
    MSM_G1, MSM_G2, ALLOC_BASES = range(1, 17)
TRANSFORM_PACKET = 17
FR, G1 = 0, 1
ADD, SUB, MUL = 0, 1, 2
DEFAULT_LIMIT = 4 << 30       # resident.rs:252
DEFAULT_JOURNAL = 64 << 20


class ResidentBackend:
    """RebuildBackend over one HipBackend owner."""

    def __init__(self, backend: HipBackend, limit: int = DEFAULT_LIMIT, journal: int = DEFAULT_JOURNAL):
        self.b = backend
        err = _err()
        root = encode_fr(root_of_unity_2_20())
        _check(backend.lib.dvgpu_resident_configure(backend.handle, limit, journal,
                                                    ctypes.cast(ctypes.c_char_p(root), ctypes.c_void_p),
                                                    err, 256), err)

    def call(self, opcode, kind=0, flags=0, a=0, b=0, c=0, a_offset=0, b_offset=0, c_offset=0,
             length=0, rows=0, cols=0, input=None, output=None, profile=False):
        op = _lib.DvgpuResidentOp(opcode, kind, flags, 0, a, b, c, a_offset, b_offset, c_offset, length,
                                  rows, cols, input, output)
        res = ctypes.c_uint64(0)
        err = _err()
        if profile:
            prof = _lib.DvgpuResidentProfile()
            _check(self.b.lib.dvgpu_resident_call_profiled(self.b.handle, ctypes.byref(op), ctypes.byref(res),
                                                           ctypes.byref(prof), err, 256), err)
            return res.value, {k: getattr(prof, k) for k, _ in prof._fields_}
        _check(self.b.lib.dvgpu_resident_call(self.b.handle, ctypes.byref(op), ctypes.byref(res), err, 256), err)
        return res.value

    # ---- allocation / metadata ---------------------------------------------
    def alloc(self, kind: int, length: int, writable_after_seal: bool = False) -> int:
        return self.call(ALLOC, kind=kind, flags=int(writable_after_seal), length=length)

    def alloc_bases(self, points: Sequence) -> int:
        buf = ctypes.create_string_buffer(b"".join(encode_g1(p) for p in points) or b"\0")
        return self.call(ALLOC_BASES, kind=0, length=len(points), input=ctypes.addressof(buf))

    def buffer_len(self, h: int) -> int:
        return self.call(LENGTH, a=h)

    def buffer_kind(self, h: int) -> int:
        return self.call(KIND, a=h)

    def seal(self) -> None:
        self.b.lib.dvgpu_resident_seal(self.b.handle)

    def begin_transaction(self) -> None:
        self.call(BEGIN)

    def rollback(self) -> None:
        self.call(ROLLBACK)

    def commit_transaction(self) -> None:
        self.b.lib.dvgpu_resident_commit(self.b.handle)

    # ---- data movement ------------------------------------------------------
    def write_fields(self, h: int, offset: int, values: Sequence[int]) -> None:
        buf = ctypes.create_string_buffer(b"".join(encode_fr(v) for v in values) or b"\0")
        self.call(WRITE, kind=FR, a=h, a_offset=offset, length=len(values), input=ctypes.addressof(buf))

    def read_fields(self, h: int, offset: int, n: int) -> List[int]:
        buf = ctypes.create_string_buffer(max(1, 32 * n))
        self.call(READ, kind=FR, a=h, a_offset=offset, length=n, output=ctypes.addressof(buf))
        return [int.from_bytes(buf.raw[32 * i:32 * i + 32], "little") for i in range(n)]

    def write_points(self, h: int, offset: int, points: Sequence) -> None:
        buf = ctypes.create_string_buffer(b"".join(encode_g1(p) for p in points) or b"\0")
        self.call(WRITE, kind=G1, a=h, a_offset=offset, length=len(points), input=ctypes.addressof(buf))

    def read_points(self, h: int, offset: int, n: int, device_affine: bool = False) -> list:
        buf = ctypes.create_string_buffer(max(1, 64 * n))
        self.call(READ, kind=G1, flags=int(device_affine), a=h, a_offset=offset, length=n,
                  output=ctypes.addressof(buf))
        return [decode_g1(buf.raw[64 * i:64 * i + 64]) for i in range(n)]

    def copy(self, kind, src, src_offset, dst, dst_offset, n):
        self.call(COPY, kind=kind, a=src, a_offset=src_offset, c=dst, c_offset=dst_offset, length=n)

    def reverse(self, kind, src, src_offset, dst, dst_offset, n):
        self.call(REVERSE, kind=kind, a=src, a_offset=src_offset, c=dst, c_offset=dst_offset, length=n)

    # ---- transforms ---------------------------------------------------------
    def transpose(self, kind, src, dst, rows, cols, start, length, twiddle: Optional[int] = None):
        root = ctypes.create_string_buffer(encode_fr(twiddle) if twiddle is not None else b"\0" * 32)
        self.call(TRANSPOSE, kind=kind, flags=int(twiddle is not None), a=src, c=dst, a_offset=start,
                  length=length, rows=rows, cols=cols, input=ctypes.addressof(root))

    def transform_rows(self, kind, src, dst, row_len, first_row, row_count, inverse):
        self.call(TRANSFORM, kind=kind, flags=int(inverse), a=src, c=dst, a_offset=first_row,
                  rows=row_count, cols=row_len)

    def transform_four_step(self, kind, src, scratch, dst, n, inverse):
        self.call(TRANSFORM_PACKET, kind=kind, flags=int(inverse), a=src, b=scratch, c=dst, length=n)

    def field_zip(self, op, left, left_offset, right, right_offset, dst, dst_offset, n):
        self.call(ZIP, kind=FR, flags=op, a=left, a_offset=left_offset, b=right, b_offset=right_offset,
                  c=dst, c_offset=dst_offset, length=n)

    def scale_points(self, points, point_offset, scalars, scalar_offset, dst, dst_offset, n):
        self.call(SCALE, kind=G1, a=points, a_offset=point_offset, b=scalars, b_offset=scalar_offset,
                  c=dst, c_offset=dst_offset, length=n)

    def msm_g1(self, points, point_offset, scalars, scalar_offset, n):
        out = ctypes.create_string_buffer(64)
        self.call(MSM_G1, a=points, a_offset=point_offset, b=scalars, b_offset=scalar_offset, length=n,
                  output=ctypes.addressof(out))
        return decode_g1(out.raw)

    def msm_g2(self, points: Sequence, scalars, scalar_offset):
        bases = ctypes.create_string_buffer(b"".join(encode_g2(p) for p in points) or b"\0")
        out = ctypes.create_string_buffer(128)
        self.call(MSM_G2, b=scalars, b_offset=scalar_offset, length=len(points),
                  input=ctypes.addressof(bases), output=ctypes.addressof(out))
        return decode_g2(out.raw)

    def stats(self):
        st = _lib.DvgpuResidentStats()
        if self.b.lib.dvgpu_resident_stats(self.b.handle, ctypes.byref(st)) != 0:
            raise BackendError(-1, "invalid resident stats owner")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def diagnostics(self):
        st = _lib.DvgpuResidentDiagnostics()
        if self.b.lib.dvgpu_resident_diagnostics(self.b.handle, ctypes.byref(st)) != 0:
            raise BackendError(-1, "invalid resident diagnostics owner")
        return {k: getattr(st, k) for k, _ in st._fields_}
