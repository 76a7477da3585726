"""Fr polynomial tools with batched device products (include/dynaverse_poly.h):
the host mirror of crates/protocol/src/domain/fast.rs.

    tree = SubproductTree(backend, points)        # SubproductTree::new
    tree.vanishing(), tree.inverse_derivatives()
    tree.multipoint_eval(poly), tree.interpolate(values), tree.cauchy_sums(weights)
    fft_multiply(backend, left, right), newton_divide_exact(backend, num, den)

Polynomials are lists of canonical ints, lowest degree first; errors of the
reference raise BackendError with status -1.
"""
import ctypes
from typing import List, Optional, Sequence

from . import _lib
from .backend import BackendError, HipBackend, _check, _err
from .codec import encode_fr, root_of_unity_2_20

_ROOT = encode_fr(root_of_unity_2_20())


def _enc(vals: Sequence[int]) -> bytes:
    return b"".join(encode_fr(v) for v in vals)


def _dec(raw: bytes, n: int) -> List[int]:
    return [int.from_bytes(raw[32 * i:32 * i + 32], "little") for i in range(n)]


def _buf(data: bytes):
    return ctypes.create_string_buffer(data or b"\0")


def fft_multiply(backend: HipBackend, left: Sequence[int], right: Sequence[int]) -> List[int]:
    """fast.rs:213-234 (fft_multiply_with_backend)."""
    cap = max(1, len(left) + len(right))
    out = ctypes.create_string_buffer(32 * cap)
    n = ctypes.c_size_t(0)
    err = _err()
    _check(backend.lib.dvgpu_poly_mul(backend.handle, _buf(_ROOT), _buf(_enc(left)), len(left), _buf(_enc(right)),
                                      len(right), out, cap, ctypes.byref(n), err, 256), err)
    return _dec(out.raw, n.value)


def newton_divide_exact(backend: HipBackend, numerator: Sequence[int], denominator: Sequence[int]) -> List[int]:
    """fast.rs:236-240 (newton_divide_exact_with_backend)."""
    cap = max(1, len(numerator))
    out = ctypes.create_string_buffer(32 * cap)
    n = ctypes.c_size_t(0)
    err = _err()
    _check(backend.lib.dvgpu_poly_divide_exact(backend.handle, _buf(_ROOT), _buf(_enc(numerator)), len(numerator),
                                               _buf(_enc(denominator)), len(denominator), out, cap, ctypes.byref(n),
                                               err, 256), err)
    return _dec(out.raw, n.value)


class SubproductTree:
    """fast.rs:9-139 on one backend."""

    def __init__(self, backend: HipBackend, points: Sequence[int]):
        self.b = backend
        self.n = len(points)
        h = ctypes.c_void_p()
        self.last_stats = _lib.DvgpuPolyStats()
        err = _err()
        _check(backend.lib.dvgpu_poly_tree_create(backend.handle, _buf(_ROOT), _buf(_enc(points)), len(points),
                                                  ctypes.byref(h), ctypes.byref(self.last_stats), err, 256), err)
        self.handle = h

    def close(self):
        if self.handle is not None and self.handle.value:
            self.b.lib.dvgpu_poly_tree_destroy(self.b.handle, self.handle)
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self) -> dict:
        return {"rounds": self.last_stats.rounds, "products": self.last_stats.products,
                "device_products": self.last_stats.device_products}

    def _vec(self, fn, cap):
        out = ctypes.create_string_buffer(32 * max(1, cap))
        n = ctypes.c_size_t(0)
        err = _err()
        _check(fn(self.b.handle, self.handle, out, cap, ctypes.byref(n), err, 256), err)
        return _dec(out.raw, n.value)

    def vanishing(self) -> List[int]:
        return self._vec(self.b.lib.dvgpu_poly_tree_vanishing, self.n + 1)

    def inverse_derivatives(self) -> List[int]:
        return self._vec(self.b.lib.dvgpu_poly_tree_inverse_derivatives, self.n)

    def multipoint_eval(self, polynomial: Sequence[int]) -> List[int]:
        out = ctypes.create_string_buffer(32 * max(1, self.n))
        err = _err()
        _check(self.b.lib.dvgpu_poly_tree_eval(self.b.handle, _buf(_ROOT), self.handle, _buf(_enc(polynomial)),
                                               len(polynomial), out, ctypes.byref(self.last_stats), err, 256), err)
        return _dec(out.raw, self.n)

    def interpolate(self, values: Sequence[int]) -> List[int]:
        cap = max(1, len(values))
        out = ctypes.create_string_buffer(32 * cap)
        n = ctypes.c_size_t(0)
        err = _err()
        _check(self.b.lib.dvgpu_poly_tree_interpolate(self.b.handle, _buf(_ROOT), self.handle, _buf(_enc(values)),
                                                      len(values), out, cap, ctypes.byref(n),
                                                      ctypes.byref(self.last_stats), err, 256), err)
        return _dec(out.raw, n.value)

    def cauchy_sums(self, weights: Sequence[int]) -> List[int]:
        out = ctypes.create_string_buffer(32 * max(1, len(weights)))
        err = _err()
        _check(self.b.lib.dvgpu_poly_tree_cauchy_sums(self.b.handle, _buf(_ROOT), self.handle, _buf(_enc(weights)),
                                                      len(weights), out, ctypes.byref(self.last_stats), err, 256),
               err)
        return _dec(out.raw, len(weights))
