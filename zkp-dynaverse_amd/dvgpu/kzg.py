"""Opening-cache precompute on the device: the host mirror of KzgParams'
``precompute_single_openings`` and its retained SRS transform
(crates/protocol/src/kzg.rs:217-300) over include/dynaverse_kzg.h.

    srs = KzgSrs(backend, g1_powers)          # KzgParams::g1_powers
    openings = srs.precompute_single_openings(coefficients, domain_size)
    q4, q5, q6 = srs.precompute_batch([f4, f5, f6], n)   # one call, one degree

Points are affine (x, y) ints or None (identity), scalars canonical ints, as
everywhere in this package.  Errors mirror kzg.rs:222-228 and raise
``BackendError`` with status -1.
"""
import ctypes
from typing import List, Optional, Sequence, Tuple

from . import _lib
from .backend import BackendError, HipBackend, _check, _err
from .codec import decode_g1, encode_fr, encode_g1, root_of_unity_2_20


class KzgSrs:
    """An SRS resident on one backend with its retained convolution basis."""

    def __init__(self, backend: HipBackend, g1_powers: Sequence = (), encoded: Optional[bytes] = None):
        self.b = backend
        data = encoded if encoded is not None else b"".join(encode_g1(p) for p in g1_powers)
        if len(data) % 64:
            raise ValueError("encoded G1 powers must be 64 bytes each")
        self.count = len(data) // 64
        buf = ctypes.create_string_buffer(data or b"\0")
        root = ctypes.create_string_buffer(encode_fr(root_of_unity_2_20()))
        out = ctypes.c_void_p()
        err = _err()
        _check(backend.lib.dvgpu_kzg_srs_create(backend.handle, buf, self.count, root, ctypes.byref(out), err, 256),
               err)
        self.handle = out
        self.last_timings = None

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            self.b.lib.dvgpu_kzg_srs_destroy(self.b.handle, self.handle)
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def retained_basis(self) -> Tuple[int, int]:
        """(degree of the retained basis or 0, number of basis builds so far)."""
        d, k = ctypes.c_uint64(0), ctypes.c_uint64(0)
        if self.b.lib.dvgpu_kzg_srs_basis(self.b.handle, self.handle, ctypes.byref(d), ctypes.byref(k)) != 0:
            raise BackendError(-1, "invalid KZG SRS handle")
        return d.value, k.value

    def precompute_bytes(self, coefficients: bytes, coeff_count: int, poly_count: int, domain_size: int,
                         offset: Optional[int] = None) -> bytes:
        """Encoded form: coefficients row-major 32 B each, result poly_count x
        domain_size x 64 B; event timings land in ``last_timings``."""
        if len(coefficients) != 32 * coeff_count * poly_count:
            raise ValueError("coefficient bytes do not match coeff_count x poly_count")
        out = ctypes.create_string_buffer(max(1, 64 * domain_size * poly_count))
        cbuf = ctypes.create_string_buffer(coefficients or b"\0")
        obuf = ctypes.create_string_buffer(encode_fr(offset)) if offset is not None else None
        t = _lib.DvgpuKzgTimings()
        err = _err()
        _check(self.b.lib.dvgpu_kzg_precompute_single_openings(self.b.handle, self.handle, cbuf, coeff_count,
                                                               poly_count, domain_size, obuf, out, ctypes.byref(t),
                                                               err, 256), err)
        self.last_timings = {k: getattr(t, k) for k, _ in t._fields_}
        return out.raw[:64 * domain_size * poly_count]

    def precompute_batch(self, polynomials: Sequence[Sequence[int]], domain_size: int,
                         offset: Optional[int] = None) -> List[list]:
        """Openings of several polynomials with the same coefficient count."""
        if not polynomials:
            return []
        count = len(polynomials[0])
        if any(len(p) != count for p in polynomials):
            raise ValueError("batched polynomials must have the same coefficient count")
        data = b"".join(encode_fr(c) for p in polynomials for c in p)
        raw = self.precompute_bytes(data, count, len(polynomials), domain_size, offset)
        n = domain_size
        return [[decode_g1(raw[64 * (r * n + i):64 * (r * n + i + 1)]) for i in range(n)]
                for r in range(len(polynomials))]

    def precompute_single_openings(self, coefficients: Sequence[int], domain_size: int,
                                   offset: Optional[int] = None) -> list:
        """kzg.rs:217-275 for one polynomial over the domain (or its coset)."""
        return self.precompute_batch([list(coefficients)], domain_size, offset)[0]


def glv_split(backend_lib, scalars: Sequence[int]) -> List[Tuple[int, int]]:
    """Signed (k1, k2) with k = k1 + k2 lambda mod r, as the device uses them."""
    data = b"".join(encode_fr(s) for s in scalars)
    out = ctypes.create_string_buffer(max(1, 40 * len(scalars)))
    if backend_lib.dvgpu_ext_glv_split(ctypes.create_string_buffer(data or b"\0"), len(scalars), out) != 0:
        raise BackendError(-1, "non-canonical scalar")
    res = []
    for i in range(len(scalars)):
        e = out.raw[40 * i:40 * i + 40]
        k1, k2 = int.from_bytes(e[0:16], "little"), int.from_bytes(e[16:32], "little")
        if int.from_bytes(e[32:36], "little"):
            k1 = -k1
        if int.from_bytes(e[36:40], "little"):
            k2 = -k2
        res.append((k1, k2))
    return res
