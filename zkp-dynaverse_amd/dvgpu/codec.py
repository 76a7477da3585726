"""Wire codec at the CPU<->GPU boundary (crates/gpu/src/codec.rs:1-107).

Scalars: 32-byte little-endian canonical integers (< r).  G1 affine: 64 bytes
x || y (canonical LE); infinity is all zero.  G2 affine: 128 bytes
x.c0 || x.c1 || y.c0 || y.c1 (crates/gpu/src/resident.rs:419-431).
Non-canonical or off-curve inputs are rejected, never reduced.

Python-side points: G1 = (x, y) ints or None; G2 = ((x0, x1), (y0, y1)) or None.
"""
import numpy as np

FQ_MODULUS = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
FR_MODULUS = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
FR_BYTES, G1_BYTES, G2_BYTES = 32, 64, 128
MAX_MSM_SIZE = 1 << 20
MAX_NTT_SIZE = 1 << 20
MAX_MSM_BATCH_SIZE = 64


def root_of_unity_2_20():
    """arkworks Fr::get_root_of_unity(2^20): 5^((r-1)/2^20) (generator 5)."""
    return pow(5, (FR_MODULUS - 1) >> 20, FR_MODULUS)


class CodecError(ValueError):
    pass


def encode_fr(v: int) -> bytes:
    if not 0 <= v < FR_MODULUS:
        raise CodecError("scalar out of range")
    return int(v).to_bytes(32, "little")


def decode_fr(b) -> int:
    b = bytes(b)
    if len(b) != FR_BYTES:
        raise CodecError("scalar must be exactly 32 bytes")
    v = int.from_bytes(b, "little")
    if v >= FR_MODULUS:
        raise CodecError("non-canonical or invalid scalar bytes")
    return v


def g1_on_curve(p) -> bool:
    if p is None:
        return True
    x, y = p
    return 0 <= x < FQ_MODULUS and 0 <= y < FQ_MODULUS and (y * y - x * x * x - 3) % FQ_MODULUS == 0


def encode_g1(p) -> bytes:
    if p is None:
        return bytes(64)
    return int(p[0]).to_bytes(32, "little") + int(p[1]).to_bytes(32, "little")


def decode_g1(b):
    b = bytes(b)
    if len(b) != G1_BYTES:
        raise CodecError("G1 point must be exactly 64 bytes")
    if b == bytes(64):
        return None
    p = (int.from_bytes(b[:32], "little"), int.from_bytes(b[32:], "little"))
    if not g1_on_curve(p):
        raise CodecError("invalid or off-curve G1 bytes")
    return p


_B2 = None


def _twist_b():
    global _B2
    if _B2 is None:
        q = FQ_MODULUS
        n = (81 + 1) % q
        ni = pow(n, q - 2, q)
        xi_inv = (9 * ni % q, (-ni) % q)
        _B2 = (3 * xi_inv[0] % q, 3 * xi_inv[1] % q)
    return _B2


def _f2mul(a, b):
    q = FQ_MODULUS
    return ((a[0] * b[0] - a[1] * b[1]) % q, (a[0] * b[1] + a[1] * b[0]) % q)


def g2_on_curve(p) -> bool:
    if p is None:
        return True
    (x0, x1), (y0, y1) = p
    if not all(0 <= v < FQ_MODULUS for v in (x0, x1, y0, y1)):
        return False
    x, y = (x0, x1), (y0, y1)
    lhs = _f2mul(y, y)
    rhs = _f2mul(_f2mul(x, x), x)
    b = _twist_b()
    return lhs == ((rhs[0] + b[0]) % FQ_MODULUS, (rhs[1] + b[1]) % FQ_MODULUS)


def encode_g2(p) -> bytes:
    if p is None:
        return bytes(128)
    (x0, x1), (y0, y1) = p
    return b"".join(int(v).to_bytes(32, "little") for v in (x0, x1, y0, y1))


def decode_g2(b):
    b = bytes(b)
    if len(b) != G2_BYTES:
        raise CodecError("G2 point must be exactly 128 bytes")
    if b == bytes(128):
        return None
    v = [int.from_bytes(b[32 * i:32 * i + 32], "little") for i in range(4)]
    p = ((v[0], v[1]), (v[2], v[3]))
    if not g2_on_curve(p):
        raise CodecError("invalid or off-curve G2 bytes")
    return p


# ---- vectorised helpers for large arrays ----------------------------------

def ints_to_array(values, width=32) -> np.ndarray:
    """List of non-negative ints -> (n, width) uint8 little-endian array."""
    buf = b"".join(int(v).to_bytes(width, "little") for v in values)
    return np.frombuffer(buf, dtype=np.uint8).reshape(len(values), width).copy()


def array_to_ints(arr: np.ndarray):
    arr = np.ascontiguousarray(arr, dtype=np.uint8)
    return [int.from_bytes(arr[i].tobytes(), "little") for i in range(arr.shape[0])]
