"""Pairing equations on the MI355X (dvgpu_bn254_pairing_equations and the
PairingChecks mirror) on the equation shapes of the reference's verifiers:
Dynamo (crates/protocol/src/dynamo.rs:423-429), the anchor relations
(dynaverse.rs:1053-1058), the delta divisors (dynaverse.rs:1081-1085),
consistency (consistency.rs:101-105) and the cache openings
(rebuilding/machine.rs:321,335).  An equation
e([a]G1, [b]G2) == prod e([c_j]G1, [d_j]G2) holds exactly when
ab == sum c_j d_j mod r (bilinearity and non-degeneracy), which fixes every
expected decision; a few are also checked with the oracle's pairing."""
import random

import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.verification import (PairingChecks, VerificationMode, VerificationPhase, VerificationTrace,
                                encode_equations, pairing_equations, pairing_equations_encoded)
from oracle import bn254 as o

pytestmark = pytest.mark.gpu

MODES = list(VerificationMode)


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


def g1(k):
    k %= o.R
    return o.g1_mul(k) if k else None


def g2(k):
    k %= o.R
    return o.mul(o.G2_GEN, k, o.FQ2) if k else None


def build(rng, right_scalars, holds, left_b=None):
    """The equation (aG, bH) == prod (c_j G, d_j H) for the given (c_j, d_j),
    with a chosen so that it holds or fails."""
    b = left_b if left_b is not None else rng.randrange(1, o.R)
    s = sum(c * d for c, d in right_scalars) % o.R
    a = s * pow(b, -1, o.R) % o.R
    if not holds:
        a = (a + rng.randrange(1, o.R)) % o.R
    return (g1(a), g2(b)), [(g1(c), g2(d)) for c, d in right_scalars]


def reference_shapes(rng):
    """One equation of each verifier shape, built to hold."""
    r = lambda: rng.randrange(1, o.R)  # noqa: E731
    return {
        "dynamo_first": build(rng, [(r(), r()) for _ in range(4)], True),           # dynamo.rs:423-425
        "dynamo_relation": build(rng, [(r(), 1)], True),                              # dynamo.rs:427
        "anchor_multiplication": build(rng, [(r(), 1), (r(), r())], True),            # dynaverse.rs:1056-1058
        "delta_divisor": build(rng, [(1, r())], True, left_b=1),                      # dynaverse.rs:1083
        "consistency_quotients": build(rng, [(r(), r()), (r(), r())], True, left_b=1),  # consistency.rs:101
        "cache_opening": build(rng, [(r(), r())], True, left_b=1),                    # machine.rs:335
    }


@pytest.mark.parametrize("mode", MODES)
def test_reference_shapes_hold_and_a_changed_point_fails(backend, mode):
    rng = random.Random(71)
    shapes = reference_shapes(rng)
    eqs = list(shapes.values())
    assert pairing_equations(backend, eqs, mode) == [True] * len(eqs)
    broken = []
    for left, right in eqs:
        p, q = right[-1]
        broken.append((left, right[:-1] + [(o.add(p, o.G1_GEN, o.FQ), q)]))
    broken.append(((eqs[0][0][0], g2(5)), eqs[0][1]))   # a changed left G2 point
    assert pairing_equations(backend, broken, mode) == [False] * len(broken)


@pytest.mark.parametrize("mode", MODES)
def test_mixed_batch_matches_the_algebra_and_single_calls(backend, mode):
    """48 equations of 1..16 pairs (some with identities), one call; every
    decision fixed by the scalars; a sample equal one equation at a time."""
    rng = random.Random(72 + int(mode))
    eqs, expect = [], []
    for i in range(48):
        pairs = 1 + (i * 7) % 16
        holds = rng.random() < 0.5
        if pairs == 1:
            eqs.append(((None, g2(3)), []) if holds else ((g1(rng.randrange(1, o.R)), g2(2)), []))
        else:
            sc = [(rng.randrange(0, o.R) if rng.random() > 0.1 else 0, rng.randrange(1, o.R))
                  for _ in range(pairs - 1)]
            eqs.append(build(rng, sc, holds))
        expect.append(holds)
    got = pairing_equations(backend, eqs, mode)
    assert got == expect
    for k in (0, 5, 17, 47):
        assert pairing_equations(backend, [eqs[k]], mode) == [expect[k]]


def test_decisions_agree_with_the_oracle_pairing(backend):
    """e(L) == prod e(R_j) through the oracle's own GT arithmetic."""
    rng = random.Random(73)
    eqs = [build(rng, [(rng.randrange(1, 50), rng.randrange(1, 50))], True),
           build(rng, [(3, 4), (5, 6)], False)]
    for left, right in eqs:
        lhs = o.pairing_product([left])
        rhs = o.pairing_product(list(right))
        for mode in MODES:
            assert pairing_equations(backend, [(left, right)], mode) == [lhs == rhs]


@pytest.mark.parametrize("mode", MODES)
def test_pairing_checks_mirror(backend, mode):
    """PairingChecks.equal / defer / run / equal_prepared on one verifier's
    phases, with the reference's counts."""
    rng = random.Random(74)
    shapes = reference_shapes(rng)
    trace = VerificationTrace()
    checks = PairingChecks(backend, mode, trace)
    started = checks.begin(VerificationPhase.ANCHOR_DYNAMO)
    ok = all(checks.equal(*shapes[k]) for k in ("dynamo_first", "dynamo_relation"))
    checks.end(started)
    assert ok
    checks.begin(VerificationPhase.DELTA_DIVISORS)
    idx = [checks.defer(*shapes[k]) for k in ("delta_divisor", "consistency_quotients", "anchor_multiplication")]
    (left, right) = shapes["cache_opening"]
    idx.append(checks.defer(left, [(o.neg(right[0][0]), right[0][1])]))   # a sign error fails
    assert checks.run() == [True, True, True, False] and idx == [0, 1, 2, 3]
    c = trace.measurements[VerificationPhase.ANCHOR_DYNAMO].counts
    loops = 5 + 2 if mode == VerificationMode.REFERENCE else 2
    assert (c.equations, c.pairing_terms, c.miller_loops, c.final_exponentiations) == (2, 7, loops, loops)
    assert trace.measurements[VerificationPhase.DELTA_DIVISORS].calls == 1
    if mode == VerificationMode.PRODUCT:
        # machine.rs:321: e(C - [v]g1, generator) == e(pi, divisor)
        (d, gen), [(pi, div)] = shapes["cache_opening"]
        assert checks.equal_prepared(d, pi, gen, div)
        assert not checks.equal_prepared(d, o.neg(pi), gen, div)


def test_invalid_inputs_fail_the_call_and_keep_the_owner(backend):
    (left, right) = build(random.Random(75), [(2, 3)], True)
    g1b, g2b, off = encode_equations([(left, right)])
    bad_curve = bytearray(g1b)
    bad_curve[64] ^= 1                                   # right G1 point off E
    cases = [(bytes(bad_curve), g2b, off, 0), (g1b, g2b, [0, 0, 2], 0), (g1b, g2b, [1, 2], 0),
             (g1b, g2b, off, 7)]
    big = bytes(64 * 17)
    cases.append((big, bytes(128 * 17), [0, 17], 0))     # 17 pairs in one equation
    noncanon = bytearray(g2b)
    noncanon[0:32] = (o.Q).to_bytes(32, "little")        # x.c0 = q
    cases.append((g1b, bytes(noncanon), off, 1))
    for a, b, offsets, mode in cases:
        with pytest.raises(BackendError) as e:
            pairing_equations_encoded(backend, a, b, offsets, mode)
        assert e.value.status == -1
    assert pairing_equations(backend, [(left, right)], VerificationMode.PRODUCT) == [True]
    assert pairing_equations(backend, [], VerificationMode.REFERENCE) == []
