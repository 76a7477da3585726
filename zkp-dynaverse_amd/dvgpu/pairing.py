"""Batched pairing-product checks on the device (include/dynaverse_pairing.h):
the host mirror of the opening-cache checks of
crates/protocol/src/dynaverse/rebuilding/machine.rs:255-390, where every
cache point pi_i of a commitment C must satisfy
e(C - [v_i] g1, g2) == e(pi_i, [tau]g2 - [z_i]g2).

    ok = pairing_check(backend, [[(P1, Q1), (P2, Q2)], ...])   # prod == 1 per group
    groups = cache_check_groups(differences, g2, openings, divisors)

Points are affine int tuples or None, as everywhere in this package.
"""
import ctypes
from typing import List, Sequence, Tuple

from .backend import HipBackend, _check, _err
from .codec import FQ_MODULUS, encode_g1, encode_g2


def _encode(groups):
    terms = len(groups[0])
    if any(len(g) != terms for g in groups):
        raise ValueError("every group needs the same number of pairs")
    g1 = b"".join(encode_g1(p) for g in groups for p, _ in g)
    g2 = b"".join(encode_g2(q) for g in groups for _, q in g)
    return terms, g1, g2


def pairing_check_encoded(backend: HipBackend, g1: bytes, g2: bytes, terms: int, count: int,
                          want_gt: bool = False):
    """Encoded form; returns (results bytes, gt bytes or None, kernel ms)."""
    res = ctypes.create_string_buffer(max(1, count))
    gt = ctypes.create_string_buffer(384 * count) if want_gt else None
    ms = ctypes.c_float(0)
    err = _err()
    _check(backend.lib.dvgpu_bn254_pairing_check(backend.handle, ctypes.c_char_p(g1 or b"\0"),
                                                 ctypes.c_char_p(g2 or b"\0"), terms, count, res, gt,
                                                 ctypes.byref(ms), err, 256), err)
    return res.raw[:count], (gt.raw if gt is not None else None), ms.value


def pairing_check(backend: HipBackend, groups: Sequence[Sequence[Tuple]]) -> List[bool]:
    """For every group of (P, Q) pairs: is prod e(P, Q) one?"""
    if not groups:
        return []
    terms, g1, g2 = _encode(groups)
    res, _, _ = pairing_check_encoded(backend, g1, g2, terms, len(groups))
    return [b == 1 for b in res]


def pairing_products(backend: HipBackend, groups) -> List[Tuple[int, ...]]:
    """The reduced products in tower order (12 canonical Fq ints each)."""
    terms, g1, g2 = _encode(groups)
    _, gt, _ = pairing_check_encoded(backend, g1, g2, terms, len(groups), want_gt=True)
    return [tuple(int.from_bytes(gt[384 * i + 32 * k:384 * i + 32 * k + 32], "little") for k in range(12))
            for i in range(len(groups))]


def _neg_g1(p):
    return None if p is None else (p[0], (-p[1]) % FQ_MODULUS)


def cache_check_groups(commitment_minus_values: Sequence, g2_gen, openings: Sequence, divisors: Sequence):
    """The two-term groups of machine.rs:333: (C - [v_i]g1, g2), (-pi_i, [tau - z_i]g2)."""
    return [[(d, g2_gen), (_neg_g1(pi), q)] for d, pi, q in zip(commitment_minus_values, openings, divisors)]


def opening_cache_check_encoded(backend: HipBackend, commitments: bytes, wires: int, g1_0: bytes, g2_0: bytes,
                                g2_1: bytes, points: bytes, n: int, values: bytes, caches: bytes):
    """Encoded form: returns (results bytes, wires x n; kernel ms)."""
    res = ctypes.create_string_buffer(max(1, wires * n))
    ms = ctypes.c_float(0)
    err = _err()
    bufs = [ctypes.c_char_p(a or b"\0") for a in (commitments, g1_0, g2_0, g2_1, points, values, caches)]
    _check(backend.lib.dvgpu_bn254_opening_cache_check(backend.handle, bufs[0], wires, bufs[1], bufs[2], bufs[3],
                                                       bufs[4], n, bufs[5], bufs[6], res, ctypes.byref(ms), err, 256),
           err)
    return res.raw[:wires * n], ms.value


def opening_cache_check(backend: HipBackend, commitments: Sequence, g1_0, g2_0, g2_1, points: Sequence[int],
                        values: Sequence[Sequence[int]], caches: Sequence[Sequence], with_time: bool = False):
    """validate_initial_caches (machine.rs:357-382) on the device: result[w][i]
    is whether e(C_w - [v_wi] g1_0, g2_0) == e(pi_wi, g2_1 - [z_i] g2_0)."""
    from .codec import encode_fr
    wires, n = len(commitments), len(points)
    if len(values) != wires or len(caches) != wires or any(len(v) != n for v in values) or \
            any(len(c) != n for c in caches):
        raise ValueError("values and caches need one row of n entries per commitment")
    raw, ms = opening_cache_check_encoded(
        backend, b"".join(encode_g1(c) for c in commitments), wires, encode_g1(g1_0), encode_g2(g2_0),
        encode_g2(g2_1), b"".join(encode_fr(z) for z in points), n,
        b"".join(encode_fr(v) for row in values for v in row), b"".join(encode_g1(p) for row in caches for p in row))
    out = [[raw[w * n + i] == 1 for i in range(n)] for w in range(wires)]
    return (out, ms) if with_time else out


def first_failure(results: Sequence[Sequence[bool]]):
    """(wire index, point index) of the first failing check in the order of
    machine.rs:361-381 (blocks of four points, wires inside a block), or None."""
    wires, n = len(results), len(results[0]) if results else 0
    for start in range(0, n, 4):
        for w in range(wires):
            for i in range(start, min(start + 4, n)):
                if not results[w][i]:
                    return w, i
    return None


def generators(backend: HipBackend) -> Tuple[bytes, bytes]:
    """(g1, g2) canonical generator encodings, [1]G through the setup entry."""
    from .codec import encode_fr
    gens = getattr(backend, "_generators", None)
    if gens is None:
        one = encode_fr(1)
        gens = (backend.fixed_base_encoded(one, 1, False), backend.fixed_base_encoded(one, 1, True))
        backend._generators = gens
    return gens


def kzg_verify_encoded(backend: HipBackend, commitments: bytes, values: bytes, openings: bytes, divisors: bytes,
                       count: int) -> Tuple[bytes, float]:
    """dvgpu_bn254_kzg_verify_batch: result byte t is 1 when
    e(C_t - [y_t] g1, g2) == e(W_t, D_t) (dynaverse.rs:447-451).  Inputs are
    count x 64 / 32 / 64 / 128 canonical bytes; returns (results, kernel ms)."""
    if not count:
        return b"", 0.0
    if len(commitments) != 64 * count or len(values) != 32 * count or len(openings) != 64 * count or \
            len(divisors) != 128 * count:
        raise ValueError("KZG verification inputs need count x 64 / 32 / 64 / 128 bytes")
    g1, g2 = generators(backend)
    res = ctypes.create_string_buffer(count)
    ms = ctypes.c_float(0)
    err = _err()
    bufs = [ctypes.c_char_p(a) for a in (g1, g2, commitments, values, openings, divisors)]
    _check(backend.lib.dvgpu_bn254_kzg_verify_batch(backend.handle, *bufs, count, res, ctypes.byref(ms), err, 256),
           err)
    return res.raw[:count], ms.value


def verify_update_proofs(backend: HipBackend, proofs: Sequence[bytes], divisors: Sequence[bytes]) -> List[bool]:
    """Pairing verification of update proofs C' (64) || W' (64) || y' (32):
    proof t holds when e(C' - [y'] g1, g2) == e(W', D_t), D_t = [tau - z]g2 of
    its session."""
    if not proofs:
        return []
    raw, _ = kzg_verify_encoded(backend, b"".join(p[:64] for p in proofs), b"".join(p[128:160] for p in proofs),
                                b"".join(p[64:128] for p in proofs), b"".join(divisors), len(proofs))
    return [b == 1 for b in raw]
