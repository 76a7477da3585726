"""Host build of the device pairing header (pairing.hpp) against the oracle's
optimal ate pairing (oracle/bn254.py, Fq12 as Fq[w]/(w^12 - 18 w^6 + 82),
independent of the device's Fq2/Fq6/Fq12 tower), and the oracle pinned by the
defining properties of a pairing: bilinearity, order r, non-degeneracy.  The
reduced value is (q^12 - 1)/r-th power of the Miller value; accept/reject of
a product check is the same for any non-degenerate bilinear pairing."""
import os
import random
import subprocess
import tempfile

import pytest

from oracle import bn254 as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zkp-dynaverse_amd", "csrc")


@pytest.fixture(scope="module")
def harness():
    exe = os.path.join(tempfile.mkdtemp(prefix="pairing_host_"), "pairing_host")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(ROOT, "tests", "native", "pairing_host.cpp")], check=True)
    proc = subprocess.Popen([exe], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)

    def call(*args):
        proc.stdin.write(" ".join(str(a) for a in args) + "\n")
        proc.stdin.flush()
        return [int(v, 16) for v in proc.stdout.readline().split()]

    yield call
    proc.stdin.close()
    proc.wait()


def tower_to_flat(v):
    """Tower coefficients (c0.a0, c0.a1, c0.a2, c1.a0, c1.a1, c1.a2; Fq2 each)
    -> the oracle's flat Fq12: (c_j)_i = a + b u sits at w^(2i + j)."""
    out = [0] * 12
    for idx in range(6):
        j, i = idx // 3, idx % 3
        k = 2 * i + j
        a, b = v[2 * idx], v[2 * idx + 1]
        out[k] = (out[k] + a - 9 * b) % o.Q
        out[k + 6] = (out[k + 6] + b) % o.Q
    return tuple(out)


def flat_to_tower(f):
    """Inverse of tower_to_flat: w^k coefficient a + b u has a = f_k + 9 f_(k+6), b = f_(k+6)."""
    vals = [0] * 12
    for idx in range(6):
        j, i = idx // 3, idx % 3
        k = 2 * i + j
        b = f[k + 6]
        vals[2 * idx], vals[2 * idx + 1] = (f[k] + 9 * b) % o.Q, b
    return vals


def h(v):
    return "%x" % v


def pair_args(p, q):
    return [h(p[0]), h(p[1]), h(q[0][0]), h(q[0][1]), h(q[1][0]), h(q[1][1])]


def test_oracle_is_a_nondegenerate_bilinear_pairing():
    e = o.pairing(o.G1_GEN, o.G2_GEN)
    assert e != o.F12_ONE and o.f12_pow(e, o.R) == o.F12_ONE
    a, b = 0x1234567890ABCDEF, 0xFEDCBA987654321
    assert o.pairing(o.g1_mul(a), o.mul(o.G2_GEN, b, o.FQ2)) == o.f12_pow(e, a * b % o.R)
    assert o.pairing(o.g1_mul(a + b), o.G2_GEN) == o.f12_mul(o.pairing(o.g1_mul(a), o.G2_GEN),
                                                              o.pairing(o.g1_mul(b), o.G2_GEN))
    assert o.pairing_product([(o.g1_mul(a), o.G2_GEN), (o.neg(o.G1_GEN), o.mul(o.G2_GEN, a, o.FQ2))]) == o.F12_ONE


def test_tower_arithmetic_matches_flat_form(harness):
    rng = random.Random(21)
    for _ in range(10):
        x = tuple(rng.randrange(o.Q) for _ in range(12))
        y = tuple(rng.randrange(o.Q) for _ in range(12))
        tx, ty = flat_to_tower(x), flat_to_tower(y)
        assert tower_to_flat(tx) == x
        assert tower_to_flat(harness("mul", *map(h, tx + ty))) == o.f12_mul(x, y)
        assert tower_to_flat(harness("sqr", *map(h, tx))) == o.f12_mul(x, x)
        assert o.f12_mul(tower_to_flat(harness("inv", *map(h, tx))), x) == o.F12_ONE
        assert tower_to_flat(harness("frob", *map(h, tx))) == o.f12_pow(x, o.Q)


def test_final_exponentiation_matches(harness):
    rng = random.Random(22)
    x = tuple(rng.randrange(o.Q) for _ in range(12))
    assert tower_to_flat(harness("fexp", *map(h, flat_to_tower(x)))) == o.final_exponentiation(x)


def test_pairings_match_the_oracle(harness):
    rng = random.Random(23)
    for _ in range(3):
        p = o.g1_mul(rng.randrange(1, o.R))
        q = o.mul(o.G2_GEN, rng.randrange(1, o.R), o.FQ2)
        got = tower_to_flat(harness("product", 1, *pair_args(p, q)))
        assert got == o.pairing(p, q)
        # Miller values differ from the oracle's by Fq2 line scalings only:
        m = tower_to_flat(harness("miller", *pair_args(p, q)))
        assert o.final_exponentiation(m) == got


def test_product_checks(harness):
    """The cache equation e(C - [v]g1, g2) == e(pi, [tau - z]g2) as a
    two-term product, true and false."""
    rng = random.Random(24)
    tau, z, f_tau, f_z = (rng.randrange(1, o.R) for _ in range(4))
    q = (f_tau - f_z) * pow(tau - z, o.R - 2, o.R) % o.R
    diff = o.g1_mul((f_tau - f_z) % o.R)
    opening = o.g1_mul(q)
    div = o.mul(o.G2_GEN, (tau - z) % o.R, o.FQ2)
    one = tower_to_flat(harness("product", 2, *pair_args(diff, o.G2_GEN), *pair_args(o.neg(opening), div)))
    assert one == o.F12_ONE
    bad = tower_to_flat(harness("product", 2, *pair_args(diff, o.G2_GEN), *pair_args(o.neg(o.add(opening, o.G1_GEN)), div)))
    assert bad != o.F12_ONE


def test_prepared_lines_give_the_same_miller_value(harness):
    """A pair whose G2 lines are prepared once (prepare_lines) multiplies the
    same Miller value as the pair computed in the loop."""
    rng = random.Random(25)
    pairs = [(o.g1_mul(rng.randrange(1, o.R)), o.mul(o.G2_GEN, rng.randrange(1, o.R), o.FQ2)) for _ in range(3)]
    args = [a for p, q in pairs for a in pair_args(p, q)]
    for k in (1, 2, 3):
        assert harness("prepared", k, *args[:6 * k]) == harness("millerprod", k, *args[:6 * k])
