"""ctypes access to the C oracle (oracle/c/libbn254_ref.so) -- test
infrastructure and the timed C CPU baseline; never part of the product."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "c", "libbn254_ref.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("C oracle not built: make -C oracle")
        lib = ctypes.CDLL(LIB)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        lib.bn254ref_msm_g1.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        lib.bn254ref_ntt.argtypes = [vp, sz, ctypes.c_int, vp, ctypes.c_int, vp]
        lib.bn254ref_fixed_base_g1.argtypes = [vp, sz, vp]
        _lib = lib
    return _lib


def _p(b):
    return ctypes.cast(ctypes.c_char_p(bytes(b)), ctypes.c_void_p)


def msm_g1(points: bytes, scalars: bytes, n: int, threads: int = 1) -> bytes:
    out = ctypes.create_string_buffer(64)
    load().bn254ref_msm_g1(_p(points), _p(scalars), n, threads, out)
    return out.raw


def ntt(values: bytes, n: int, inverse: bool, root: bytes, threads: int = 1) -> bytes:
    out = ctypes.create_string_buffer(32 * n)
    if load().bn254ref_ntt(_p(values), n, int(inverse), _p(root), threads, out) != 0:
        raise ValueError("bad NTT length")
    return out.raw


def fixed_base_g1(scalars: bytes, n: int) -> bytes:
    out = ctypes.create_string_buffer(64 * n)
    load().bn254ref_fixed_base_g1(_p(scalars), n, out)
    return out.raw


def kzg_openings(powers: bytes, count: int, coefficients: bytes, ncoeff: int, n: int, root: bytes,
                 offset: bytes = None, threads: int = 1) -> bytes:
    lib = load()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.bn254ref_kzg_openings.argtypes = [vp, sz, vp, sz, sz, vp, vp, ctypes.c_int, vp]
    out = ctypes.create_string_buffer(64 * n)
    rc = lib.bn254ref_kzg_openings(_p(powers or b"\0"), count, _p(coefficients or b"\0"), ncoeff, n,
                                   _p(offset) if offset is not None else None, _p(root), threads, out)
    if rc != 0:
        raise ValueError("invalid opening-cache arguments")
    return out.raw
