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
from .codec import FR_MODULUS, decode_g1, decode_g2, encode_fr, encode_g1, encode_g2, root_of_unity_2_20

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


class FourStepPlan:
    """Natural-order four-step transform of length n as a resumable cursor
    of resident calls (crates/protocol/src/incremental_rebuild/operators.rs:
    505-597).  With a = 2^floor(log2 n / 2) and b = n / a:

      phase 0  transpose input (a x b) -> output        tile_len elements a call
      phase 1  b transforms of length a, output -> scratch   row_batch rows a call
      phase 2  transpose scratch (b x a) -> output, element (row, col) times
               w^(row * col)                              tile_len elements a call
      phase 3  a transforms of length b, output -> scratch   row_batch rows a call
      phase 4  transpose scratch (a x b) -> output        tile_len elements a call

    w is the generator of order n (its inverse for the inverse transform;
    the row transforms scale by 1/a and 1/b, so the whole scales by 1/n).
    A failed step leaves the cursor where it was; copy.copy() of a plan is
    the reference's clone, kept by a caller until its transaction commits.
    """

    PHASES = 5

    def __init__(self, kind, n, input, scratch, output, tile_len, row_batch, inverse):
        if n < 1 or n & (n - 1):
            raise ValueError("four-step length must be a power of two")
        if tile_len <= 0 or row_batch <= 0:
            raise ValueError("four-step block bounds must be positive")
        if input == scratch or input == output or scratch == output:
            raise ValueError("four-step buffers must be distinct")
        if n > 1 << 20:
            raise ValueError("unsupported four-step root order")
        log = n.bit_length() - 1
        self.kind, self.n, self.inverse = kind, n, inverse
        self.a = 1 << (log // 2)
        self.b = n // self.a
        self.input, self.scratch, self.output = input, scratch, output
        self.tile_len, self.row_batch = tile_len, row_batch
        root = pow(root_of_unity_2_20(), (1 << 20) // n, FR_MODULUS)
        self.root = pow(root, FR_MODULUS - 2, FR_MODULUS) if inverse else root
        tiles = -(-n // tile_len)
        self.total_steps = 3 * tiles + -(-self.a // row_batch) + -(-self.b // row_batch)
        self.phase, self.cursor = 0, 0

    def dimensions(self):
        return self.a, self.b

    def is_complete(self) -> bool:
        return self.phase == self.PHASES

    def initialization_plan(self) -> "FourStepPlan":
        """Initialization pays no per-update latency bound: contiguous rows
        and tiles of up to 65536 elements a call."""
        cap = 65536
        return FourStepPlan(self.kind, self.n, self.input, self.scratch, self.output, min(self.n, cap),
                            max(1, cap // max(self.a, self.b)), self.inverse)

    def execute_packet(self, backend: ResidentBackend) -> None:
        """The whole transform as one resident packet call (n <= 4096)."""
        if self.phase or self.cursor:
            raise ValueError("transform packet must start at an untouched plan")
        backend.transform_four_step(self.kind, self.input, self.scratch, self.output, self.n, self.inverse)

    def next_step(self, backend: ResidentBackend) -> bool:
        """Run one tile or row batch; True once the last one has run."""
        if self.is_complete():
            return True
        if self.phase == 0 and self.cursor == 0:
            for h in (self.input, self.scratch, self.output):
                if backend.buffer_kind(h) != self.kind:
                    raise ValueError("four-step buffer kind mismatch")
                if backend.buffer_len(h) < self.n:
                    raise ValueError("four-step buffer too short")
        phase_len = {1: self.b, 3: self.a}.get(self.phase, self.n)
        block = self.row_batch if self.phase in (1, 3) else self.tile_len
        count = min(block, phase_len - self.cursor)
        a, b, k = self.a, self.b, self.kind
        if self.phase == 0:
            backend.transpose(k, self.input, self.output, a, b, self.cursor, count)
        elif self.phase == 1:
            backend.transform_rows(k, self.output, self.scratch, a, self.cursor, count, self.inverse)
        elif self.phase == 2:
            backend.transpose(k, self.scratch, self.output, b, a, self.cursor, count, twiddle=self.root)
        elif self.phase == 3:
            backend.transform_rows(k, self.output, self.scratch, b, self.cursor, count, self.inverse)
        else:
            backend.transpose(k, self.scratch, self.output, a, b, self.cursor, count)
        self.cursor += count
        if self.cursor == phase_len:
            self.phase, self.cursor = self.phase + 1, 0
        return self.is_complete()

    def run(self, backend: ResidentBackend) -> int:
        """Every remaining step; returns the number run."""
        steps = 0
        while not self.is_complete():
            self.next_step(backend)
            steps += 1
        return steps
