"""CPU checks for the opening-cache precompute (crates/protocol/src/kzg.rs:
217-275): the oracle against the closed form the reference's own tests rely
on (``assert_dense_oracle``, kzg.rs:395-404: every cached opening equals the
dense single opening [(f(tau) - f(z)) / (tau - z)] G1 for the trapdoor tau of
``KzgParams::from_tau``), the C oracle against the Python oracle, and the
library's host-side GLV split against the lattice identity it must satisfy.
No device calls."""
import ctypes
import os
import random

import pytest

from oracle import bn254 as o
from oracle import cref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "zkp-dynaverse_amd", "dvgpu", "libdynaverse_hip.so")
ROOT_2_20 = o.fr_to_bytes(o.fr_root_of_unity(1 << 20))


def coefficients(length, offset):
    """kzg.rs:384-390: f_i = i^2 + 3 i + offset."""
    return [i * i + 3 * i + offset for i in range(length)]


def dense(tau, coeffs, n, offset=None):
    w = o.fr_root_of_unity(n)
    g = 1 if offset is None else offset
    return [o.kzg_opening_closed_form(tau, coeffs, g * pow(w, i, o.R) % o.R) for i in range(n)]


@pytest.fixture(scope="module")
def srs16():
    return o.kzg_srs(7, 16)   # KzgParams::from_tau(16, Fr::from(7)) (kzg.rs:373)


def test_group_fft_equals_dft_definition():
    rng = random.Random(5)
    for n in (1, 2, 4, 8):
        pts = [o.g1_mul(rng.randrange(o.R)) for _ in range(n)]
        if n > 2:
            pts[1] = None
        for inv in (False, True):
            assert o.ecntt_fft(pts, inv) == o.ecntt(pts, inv)


@pytest.mark.parametrize("length,n,offset", [(8, 8, 7), (8, 8, 19), (8, 16, 11), (7, 16, 11), (4, 16, 11),
                                             (7, 8, 13), (2, 8, 3), (3, 4, 9)])
def test_oracle_openings_equal_dense_openings(srs16, length, n, offset):
    """kzg.rs:372-453 shapes: same degree twice, degree switches 8/7/4/8."""
    c = coefficients(length, offset)
    assert o.kzg_openings(srs16, c, n) == dense(7, c, n)


def test_oracle_coset_openings(srs16):
    """kzg.rs:491-503: the coset offset 5 of the 16-point domain."""
    c = coefficients(7, 13)
    assert o.kzg_openings(srs16, c, 16, offset=5) == dense(7, c, 16, offset=5)


def test_oracle_constants_and_bounds():
    """kzg.rs:466-488: constants give identities; too many coefficients fail."""
    srs = o.kzg_srs(7, 8)
    for poly in ([], [3]):
        assert o.kzg_openings(srs, poly, 8) == [None] * 8
    with pytest.raises(ValueError):
        o.kzg_openings(srs, coefficients(9, 3), 8)
    with pytest.raises(ValueError):
        o.kzg_openings(srs, coefficients(10, 3), 16)


def test_c_oracle_equals_python_oracle():
    rng = random.Random(17)
    tau = 123456789
    srs = o.kzg_srs(tau, 40)
    pb = b"".join(o.g1_to_bytes(p) for p in srs)
    for length, n, offset in [(8, 8, None), (33, 64, None), (17, 32, 5), (2, 4, None), (1, 8, None),
                              (0, 4, None)]:
        c = [rng.randrange(o.R) for _ in range(length)]
        cb = b"".join(o.fr_to_bytes(v) for v in c)
        got = cref.kzg_openings(pb, len(srs), cb, length, n, ROOT_2_20,
                                o.fr_to_bytes(offset) if offset else None, threads=4)
        assert got == b"".join(o.g1_to_bytes(p) for p in o.kzg_openings(srs, c, n, offset)), (length, n)
    with pytest.raises(ValueError):
        cref.kzg_openings(pb, len(srs), bytes(32 * 9), 9, 8, ROOT_2_20)


def test_c_oracle_sampled_dense_openings_at_4096():
    rng = random.Random(3)
    tau, n = 987654321, 1 << 12
    srs_b = cref.fixed_base_g1(b"".join(o.fr_to_bytes(pow(tau, i, o.R)) for i in range(n)), n)
    c = [rng.randrange(o.R) for _ in range(n)]
    out = cref.kzg_openings(srs_b, n, b"".join(o.fr_to_bytes(v) for v in c), n, n, ROOT_2_20, threads=8)
    w = o.fr_root_of_unity(n)
    for i in (0, 1, 777, 2048, n - 1):
        assert out[64 * i:64 * i + 64] == o.g1_to_bytes(o.kzg_opening_closed_form(tau, c, pow(w, i, o.R)))


# ---- host GLV split of the library (dvgpu_ext_glv_split) --------------------

def _split(scalars):
    lib = ctypes.CDLL(LIB)
    lib.dvgpu_ext_glv_split.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    data = b"".join(o.fr_to_bytes(s) for s in scalars)
    out = ctypes.create_string_buffer(40 * len(scalars))
    assert lib.dvgpu_ext_glv_split(ctypes.c_char_p(data), len(scalars), out) == 0
    res = []
    for i in range(len(scalars)):
        e = out.raw[40 * i:40 * i + 40]
        k1, k2 = int.from_bytes(e[:16], "little"), int.from_bytes(e[16:32], "little")
        res.append((-k1 if e[32] else k1, -k2 if e[36] else k2))
    return res


def _glv_lambda():
    """The eigenvalue of (x, y) -> (beta x, y) on G1, found from the group."""
    beta = next(c for c in (pow(g, (o.Q - 1) // 3, o.Q) for g in range(2, 50)) if c != 1)
    lam = next(c for c in (pow(g, (o.R - 1) // 3, o.R) for g in range(2, 50)) if c != 1)
    target = (beta * o.G1_GEN[0] % o.Q, o.G1_GEN[1])
    return lam if o.g1_mul(lam) == target else lam * lam % o.R


def test_glv_split_is_a_short_decomposition():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lam = _glv_lambda()
    assert o.g1_mul(lam) != o.G1_GEN and (lam * lam + lam + 1) % o.R == 0
    rng = random.Random(11)
    ks = [0, 1, 2, o.R - 1, o.R - 2, lam, o.R - lam, (1 << 253), (1 << 126), (1 << 127) - 1]
    ks += [rng.randrange(o.R) for _ in range(4000)]
    for k, (k1, k2) in zip(ks, _split(ks)):
        assert (k1 + k2 * lam - k) % o.R == 0, k
        assert abs(k1) < 1 << 126 and abs(k2) < 1 << 126, k


def test_glv_split_rejects_non_canonical():
    lib = ctypes.CDLL(LIB)
    lib.dvgpu_ext_glv_split.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    out = ctypes.create_string_buffer(40)
    assert lib.dvgpu_ext_glv_split(ctypes.c_char_p(o.R.to_bytes(32, "little")), 1, out) == -1


def test_joint_sparse_form_of_the_split():
    """The recoding the shared-twiddle stages use: digits in {-1, 0, 1} that
    sum back to |k1| and |k2|, within 127 positions, about half the columns
    nonzero, and no two adjacent nonzero columns among any three."""
    lib = ctypes.CDLL(LIB)
    lib.dvgpu_ext_glv_jsf.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    rng = random.Random(13)
    ks = [0, 1, o.R - 1, 1 << 200] + [rng.randrange(o.R) for _ in range(1500)]
    data = b"".join(o.fr_to_bytes(k) for k in ks)
    out = ctypes.create_string_buffer(64 * len(ks))
    assert lib.dvgpu_ext_glv_jsf(ctypes.c_char_p(data), len(ks), out) == 0
    nonzero = 0
    for i, (k, (k1, k2)) in enumerate(zip(ks, _split(ks))):
        m = [int.from_bytes(out.raw[64 * i + 16 * q:64 * i + 16 * q + 16], "little") for q in range(4)]
        assert not (m[1] & ~m[0]) and not (m[3] & ~m[2])          # signs only on nonzero digits
        assert m[0] >> 127 == 0 and m[2] >> 127 == 0
        v1 = sum((-1 if m[1] >> l & 1 else 1) << l for l in range(128) if m[0] >> l & 1)
        v2 = sum((-1 if m[3] >> l & 1 else 1) << l for l in range(128) if m[2] >> l & 1)
        assert (v1, v2) == (abs(k1), abs(k2)), k
        cols = m[0] | m[2]
        nonzero += bin(cols).count("1")
        for l in range(126):                                       # JSF: of 3 consecutive columns one is zero
            assert (cols >> l) & 7 != 7, k
    assert nonzero / len(ks) < 0.55 * 127
