"""Pins the oracle before anything is checked against it.

Every oracle function is checked against something that does not restate its
own formula: the closed-form known answers the reference's tests use,
the DFT definition, group-order and endomorphism identities, and (for the C
oracle) the independent Python big-int oracle.
"""
import random

import pytest

from oracle import bn254 as o
from oracle import cref


def test_curve_constants_and_group_orders():
    assert o.on_curve(o.G1_GEN) and o.mul(o.G1_GEN, o.R) is None
    assert o.on_curve(o.G2_GEN, o.FQ2) and o.mul(o.G2_GEN, o.R, o.FQ2) is None
    assert o.mul(o.G1_GEN, o.R - 1) == o.neg(o.G1_GEN)
    # BN parameter: q and r are the BN polynomials at x (6x^2 = lambda).
    x = o.BN_SEED
    assert o.Q == 36 * x**4 + 36 * x**3 + 24 * x**2 + 6 * x + 1
    assert o.R == 36 * x**4 + 36 * x**3 + 18 * x**2 + 6 * x + 1
    assert bin(x).count("1") == 28 and x.bit_length() == 63


def test_root_of_unity_orientation():
    # arkworks derives roots from generator 5: w_2^28 = 5^((r-1)/2^28).
    w28 = o.fr_root_of_unity(1 << 28)
    assert pow(w28, 1 << 27, o.R) == o.R - 1
    # Published arkworks / gnark BN254 Fr two-adic root of unity.
    assert w28 == 19103219067921713944291392827692070036145651957329286315305642004821462161904
    assert o.fr_root_of_unity(1 << 20) == pow(w28, 1 << 8, o.R)


def test_ntt_equals_dft_definition_and_known_answers():
    rng = random.Random(1)
    for n in (1, 2, 4, 8, 16, 64):
        v = [rng.randrange(o.R) for _ in range(n)]
        for inv in (False, True):
            assert o.ntt(v, inv) == o.dft_bruteforce(v, inv)
        assert o.ntt(o.ntt(v, False), True) == v
    # crates/gpu/tests/ntt_backend.rs:64-86
    assert o.ntt([1] + [0] * 7) == [1] * 8
    assert o.ntt([5] * 8) == [40] + [0] * 7
    assert o.ntt([40] + [0] * 7, True) == [5] * 8
    # Convolution theorem: NTT(a * b cyclic) = NTT(a) . NTT(b)
    n = 16
    a = [rng.randrange(o.R) for _ in range(n)]
    b = [rng.randrange(o.R) for _ in range(n)]
    cyc = [sum(a[j] * b[(k - j) % n] for j in range(n)) % o.R for k in range(n)]
    assert o.ntt(cyc) == [x * y % o.R for x, y in zip(o.ntt(a), o.ntt(b))]


def test_msm_closed_forms():
    G = o.G1_GEN
    n = 40
    pts = [o.g1_mul(i) for i in range(1, n + 1)]
    # icicle-smoke/msm_smoke.cu:9-27: sum i * [i]G = [sum i^2] G
    assert o.straight_msm(pts, list(range(1, n + 1))) == o.g1_mul(sum(i * i for i in range(1, n + 1)))
    assert o.straight_msm([G] * 9, [3] * 9) == o.g1_mul(27)          # 3n G
    assert o.straight_msm([G, G, None], [2, 3, 99]) == o.g1_mul(5)    # 5G
    assert o.straight_msm([G, G], [1, o.R - 1]) is None              # G - G
    assert o.straight_msm([G] * 3, [0, 0, 0]) is None
    rng = random.Random(2)
    ks = [rng.randrange(o.R) for _ in range(5)]
    ss = [rng.randrange(o.R) for _ in range(5)]
    assert o.straight_msm([o.g1_mul(k) for k in ks], ss) == o.g1_mul(sum(k * s for k, s in zip(ks, ss)))


def test_psi_and_subgroup_predicates():
    g = o.G2_GEN
    # psi is the 6x^2 eigen-map on the r-torsion (pins the psi coefficients).
    assert o.psi(g) == o.mul(g, o.LAMBDA, o.FQ2)
    assert o.mul(o.psi(o.psi(g)), 1, o.FQ2) == o.mul(g, o.LAMBDA * o.LAMBDA % o.R, o.FQ2)
    t = o.twist_point_from_x(1)
    assert o.on_curve(t, o.FQ2)
    for alg in (1, 2, 3):
        assert o.SUBGROUP_ALGORITHMS[alg](g)
        assert o.SUBGROUP_ALGORITHMS[alg](o.g2_mul(123456789))
        assert not o.SUBGROUP_ALGORITHMS[alg](t)


def test_g2_wire_parser_rules():
    g = o.g2_to_bytes(o.G2_GEN)
    assert o.g2_parse_wire(bytes(128)) == (0, None)
    flagged = bytearray(128)
    flagged[127] = 0x40
    assert o.g2_parse_wire(bytes(flagged)) == (0, None)
    bad = bytearray(g)
    bad[127] |= 0xC0
    assert o.g2_parse_wire(bytes(bad))[0] == 1
    bad = bytearray(g)
    bad[0:32] = o.Q.to_bytes(32, "little")
    assert o.g2_parse_wire(bytes(bad))[0] == 1
    masks, rep = o.g2_masks(g + bytes(128), 3)
    assert masks == [0, 0] and rep["first_invalid"] == (1 << 64) - 1


def test_c_oracle_matches_python_oracle():
    rng = random.Random(3)
    G = o.G1_GEN
    for n in (1, 2, 3, 17, 100):
        ks = [rng.randrange(o.R) for _ in range(n)]
        pts = [None if i % 7 == 0 else o.g1_mul(k) for i, k in enumerate(ks)]
        sc = [0 if i % 11 == 0 else rng.randrange(o.R) for i in range(n)]
        exp = o.g1_mul(sum(s * k for i, (s, k) in enumerate(zip(sc, ks)) if i % 7) % o.R)
        pb = b"".join(o.g1_to_bytes(p) for p in pts)
        sb = b"".join(o.fr_to_bytes(s) for s in sc)
        for threads in (1, 4):
            assert o.g1_from_bytes(cref.msm_g1(pb, sb, n, threads)) == exp, (n, threads)
    assert o.g1_from_bytes(cref.msm_g1(o.g1_to_bytes(G) * 64, o.fr_to_bytes(3) * 64, 64, 4)) == o.g1_mul(192)
    root = o.fr_to_bytes(o.ROOT_2_20)
    for n in (2, 8, 256, 1024):
        v = [rng.randrange(o.R) for _ in range(n)]
        vb = b"".join(o.fr_to_bytes(x) for x in v)
        for inv in (False, True):
            out = cref.ntt(vb, n, inv, root, threads=3)
            assert [o.bytes_to_int(out[32 * i:32 * i + 32]) for i in range(n)] == o.ntt(v, inv)
    sc = [0, 1, o.R - 1, rng.randrange(o.R)]
    out = cref.fixed_base_g1(b"".join(o.fr_to_bytes(s) for s in sc), 4)
    assert [o.g1_from_bytes(out[64 * i:64 * i + 64]) for i in range(4)] == [o.g1_mul(s) for s in sc]
