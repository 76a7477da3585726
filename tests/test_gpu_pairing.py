"""Batched pairing-product checks on the MI355X (dynaverse_pairing.h) against
the oracle's optimal ate pairing (reduced values bit-exact) and on the
opening-cache equation of crates/protocol/src/dynaverse/rebuilding/
machine.rs:333,357-382 with openings from the device opening cache."""
import random

import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.codec import encode_g1, encode_g2
from dvgpu.kzg import KzgSrs
from dvgpu.pairing import cache_check_groups, pairing_check, pairing_check_encoded, pairing_products
from oracle import bn254 as o
from tests.test_pairing_host import tower_to_flat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


def g2(k):
    return o.mul(o.G2_GEN, k % o.R, o.FQ2)


def test_reduced_values_equal_the_oracle(backend):
    rng = random.Random(31)
    singles = [[(o.g1_mul(rng.randrange(1, o.R)), g2(rng.randrange(1, o.R)))] for _ in range(4)]
    singles.append([(o.G1_GEN, o.G2_GEN)])
    for g, v in zip(singles, pairing_products(backend, singles)):
        assert tower_to_flat(v) == o.pairing_product(g)
    pair = [[(o.G1_GEN, o.G2_GEN), (o.g1_mul(5), g2(7))]]
    assert tower_to_flat(pairing_products(backend, pair)[0]) == o.pairing_product(pair[0])
    # identity pairs contribute one
    assert pairing_check(backend, [[(None, o.G2_GEN)], [(o.G1_GEN, None)], [(None, None)]]) == [True] * 3


def test_product_checks_accept_and_reject(backend):
    rng = random.Random(32)
    groups, expect = [], []
    for i in range(40):
        a, b = rng.randrange(1, o.R), rng.randrange(1, o.R)
        good = [(o.g1_mul(a), g2(b)), (o.neg(o.g1_mul(a * b % o.R)), o.G2_GEN)]
        if i % 3 == 1:
            good[1] = (o.neg(o.g1_mul((a * b + 1) % o.R)), o.G2_GEN)
        if i % 5 == 4:
            good[0] = (good[0][0], g2(b + 1))
        groups.append(good)
        expect.append(i % 3 != 1 and i % 5 != 4)
    assert pairing_check(backend, groups) == expect


def test_three_term_groups(backend):
    """e(aG, bH) e(cG, dH) e(-(ab + cd)G, H) == 1."""
    a, b, c, d = 3, 5, 7, 11
    grp = [(o.g1_mul(a), g2(b)), (o.g1_mul(c), g2(d)), (o.neg(o.g1_mul(a * b + c * d)), o.G2_GEN)]
    bad = grp[:2] + [(o.neg(o.g1_mul(a * b + c * d + 1)), o.G2_GEN)]
    assert pairing_check(backend, [grp, bad]) == [True, False]


def test_opening_cache_points_pass_the_cache_equation(backend):
    """Every opening of a device-computed cache satisfies
    e(C - [f(z)]g1, g2) == e(pi, [tau]g2 - [z]g2); a changed opening fails."""
    rng = random.Random(33)
    tau, n = rng.randrange(2, o.R), 16
    srs = o.kzg_srs(tau, n)
    coeffs = [rng.randrange(o.R) for _ in range(n)]
    with KzgSrs(backend, srs) as z:
        openings = z.precompute_single_openings(coeffs, n)
    w = o.fr_root_of_unity(n)

    def ev(x):
        acc = 0
        for c in reversed(coeffs):
            acc = (acc * x + c) % o.R
        return acc
    commitment = o.straight_msm(srs[:n], coeffs)
    zs = [pow(w, i, o.R) for i in range(n)]
    diffs = [o.add(commitment, o.neg(o.g1_mul(ev(zi)))) for zi in zs]
    tau_g2 = g2(tau)
    divisors = [o.add(tau_g2, o.neg(g2(zi), o.FQ2), o.FQ2) for zi in zs]
    groups = cache_check_groups(diffs, o.G2_GEN, openings, divisors)
    assert pairing_check(backend, groups) == [True] * n
    groups[5] = cache_check_groups([diffs[5]], o.G2_GEN, [o.add(openings[5], o.G1_GEN)], [divisors[5]])[0]
    res = pairing_check(backend, groups)
    assert res == [i != 5 for i in range(n)]


def test_invalid_inputs(backend):
    off_curve = (1, 3)
    with pytest.raises(BackendError) as e:
        pairing_check(backend, [[(off_curve, o.G2_GEN)]])
    assert e.value.status == -1
    bad_g2 = ((1, 0), (1, 0))
    with pytest.raises(BackendError):
        pairing_check(backend, [[(o.G1_GEN, bad_g2)]])
    with pytest.raises(BackendError):
        pairing_check_encoded(backend, encode_g1(o.G1_GEN), encode_g2(o.G2_GEN), 0, 1)
    assert pairing_check(backend, [[(o.G1_GEN, o.G2_GEN)]]) == [False]   # backend still usable


def test_large_batch(backend):
    """65536 two-term checks in one call, a tiled mix of true and false."""
    rng = random.Random(34)
    base, expect = [], []
    for i in range(64):
        a, b = rng.randrange(1, o.R), rng.randrange(1, o.R)
        k = a * b % o.R if i % 7 else (a * b + 1) % o.R
        base.append([(o.g1_mul(a), g2(b)), (o.neg(o.g1_mul(k)), o.G2_GEN)])
        expect.append(i % 7 != 0)
    g1 = b"".join(encode_g1(p) for g in base for p, _ in g)
    g2b = b"".join(encode_g2(q) for g in base for _, q in g)
    reps = 1024
    res, _, ms = pairing_check_encoded(backend, g1 * reps, g2b * reps, 2, 64 * reps)
    print(f"65536 two-term checks: {ms:.2f} ms", flush=True)
    assert [b == 1 for b in res] == expect * reps


def test_device_opening_cache_check(backend):
    """validate_initial_caches on three device-computed caches of one anchor:
    all pass; a changed cache point, a changed witness value and a swapped
    commitment each fail at their own (wire, index), found first in the
    reference's block order."""
    from dvgpu.pairing import first_failure, opening_cache_check
    rng = random.Random(35)
    tau, n = rng.randrange(2, o.R), 16
    srs = o.kzg_srs(tau, n)
    w = o.fr_root_of_unity(n)
    zs = [pow(w, i, o.R) for i in range(n)]
    polys = [[rng.randrange(o.R) for _ in range(n)] for _ in range(3)]
    with KzgSrs(backend, srs) as z:
        caches = z.precompute_batch(polys, n)

    def ev(c, x):
        acc = 0
        for a in reversed(c):
            acc = (acc * x + a) % o.R
        return acc
    values = [[ev(c, zi) for zi in zs] for c in polys]
    commits = [o.straight_msm(srs[:n], c) for c in polys]
    g2_1 = g2(tau)
    res = opening_cache_check(backend, commits, o.G1_GEN, o.G2_GEN, g2_1, zs, values, caches)
    assert res == [[True] * n for _ in range(3)] and first_failure(res) is None
    bad_caches = [list(c) for c in caches]
    bad_caches[2][9] = o.add(bad_caches[2][9], o.G1_GEN)
    bad_values = [list(v) for v in values]
    bad_values[0][13] = (bad_values[0][13] + 1) % o.R
    res = opening_cache_check(backend, commits, o.G1_GEN, o.G2_GEN, g2_1, zs, bad_values, bad_caches)
    fails = [(wi, i) for wi in range(3) for i in range(n) if not res[wi][i]]
    assert fails == [(0, 13), (2, 9)]
    assert first_failure(res) == (2, 9)                 # block 8..11 precedes block 12..15
    res = opening_cache_check(backend, [commits[1], commits[0], commits[2]], o.G1_GEN, o.G2_GEN, g2_1, zs,
                              values, caches)
    assert [all(r) for r in res] == [False, False, True]
    with pytest.raises(BackendError):
        opening_cache_check(backend, commits, o.G1_GEN, o.G2_GEN, g2_1, zs, values,
                            [[(1, 3)] + list(c[1:]) for c in caches])
