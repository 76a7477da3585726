"""Host G1 arithmetic of the MSM epilogue (csrc/host_g1.hpp) against the
oracle: Montgomery products, sums, differences and inverses over Fq, the
conversion from the device's Montgomery form (R = 2^261, values up to 2^256)
and XYZZ doubling and addition, including the identity, P + P, P + (-P) and
points with non-unit ZZ."""
import os
import random
import subprocess
import tempfile

import pytest

from oracle import bn254 as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zkp-dynaverse_amd", "csrc")
Q = o.Q
RDEV = pow(2, 261, Q)


@pytest.fixture(scope="module")
def call():
    exe = os.path.join(tempfile.mkdtemp(prefix="host_g1_"), "host_g1")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(ROOT, "tests", "native", "host_g1.cpp")], check=True)
    proc = subprocess.Popen([exe], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)

    def run(*args):
        proc.stdin.write(" ".join(args) + "\n")
        proc.stdin.flush()
        return proc.stdout.readline().split()

    yield run
    proc.stdin.close()
    proc.wait()


def dev(x, lazy=False):
    """Device Montgomery words of x as hex (optionally the + p representative)."""
    v = x * RDEV % Q
    if lazy and v + Q < 1 << 256:
        v += Q
    return "%x" % v


def xyzz(P, z=1, lazy=False):
    """Device XYZZ words of affine P with Z = z (ZZ = z^2, ZZZ = z^3)."""
    if P is None:
        return ["0"] * 4
    x, y = P
    zz, zzz = z * z % Q, z * z * z % Q
    return [dev(x * zz % Q, lazy), dev(y * zzz % Q, lazy), dev(zz, lazy), dev(zzz, lazy)]


def aff(out):
    return None if out == ["inf"] else (int(out[0], 16), int(out[1], 16))


def test_field_ops(call):
    rng = random.Random(3)
    vals = [0, 1, 2, Q - 1, Q - 2, 1 << 253] + [rng.randrange(Q) for _ in range(40)]
    for a in vals:
        b = rng.choice(vals)
        assert int(call("mul", dev(a), dev(b, True))[0], 16) == a * b % Q
        assert int(call("add", dev(a, True), dev(b))[0], 16) == (a + b) % Q
        assert int(call("sub", dev(a), dev(b))[0], 16) == (a - b) % Q
        if a:
            assert int(call("inv", dev(a, True))[0], 16) == pow(a, Q - 2, Q)


def test_point_ops(call):
    rng = random.Random(4)
    for _ in range(20):
        P = o.g1_mul(rng.randrange(1, o.R))
        Qp = o.g1_mul(rng.randrange(1, o.R))
        z1, z2 = rng.randrange(1, Q), rng.randrange(1, Q)
        assert aff(call("dbl", *xyzz(P, z1, True))) == o.double(P)
        assert aff(call("addp", *xyzz(P, z1), *xyzz(Qp, z2, True))) == o.add(P, Qp)
        assert aff(call("addp", *xyzz(P, z1), *xyzz(P, z2))) == o.double(P)          # P + P
        assert aff(call("addp", *xyzz(P, z1), *xyzz(o.neg(P), z2))) is None          # P + (-P)
        assert aff(call("addp", *xyzz(None), *xyzz(Qp, z2))) == Qp                   # O + Q
        assert aff(call("addp", *xyzz(P, z1), *xyzz(None))) == P
    assert call("dbl", *xyzz(None)) == ["inf"]


def test_repeated_doubling(call):
    """dbl_n (Jacobian doublings between XYZZ conversions) against 2^n P."""
    rng = random.Random(5)
    for n in (1, 2, 3, 13, 15, 16):
        P = o.g1_mul(rng.randrange(1, o.R))
        z = rng.randrange(1, Q)
        assert aff(call("dbln", str(n), *xyzz(P, z, n % 2 == 1))) == o.mul(P, 1 << n)
    assert call("dbln", "5", *xyzz(None)) == ["inf"]
