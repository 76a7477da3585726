"""Host logic of the PairingChecks mirror (dvgpu/verification.py): batch
encoding, the mode constants of include/dynaverse_pairing.h and the trace
accounting of crates/protocol/src/verification.rs:37-45,55-68, over a stub
library (no device calls)."""
import ctypes
import os
import re

import pytest

from dvgpu.codec import encode_g1, encode_g2
from dvgpu.verification import (MAX_PAIRS, PairingChecks, VerificationMode, VerificationPhase, VerificationTrace,
                                encode_equations)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = (1, 2)
Q = ((10857046999023057135944570762232829481370756359578518086990519993285655852781,
      11559732032986387107991004021392285783925812861821192530917403151452391805634),
     (8495653923123431417604973247489272438418190587263600148770280649306958101930,
      4082367875863433681332203403145435568316851327593401208105741076214120093531))


class StubLib:
    """Records each call's offsets and mode and answers with fixed verdicts."""

    def __init__(self, verdicts):
        self.verdicts = list(verdicts)
        self.calls = []

    def dvgpu_bn254_pairing_equations(self, handle, g1, g2, offsets, count, mode, res, ms, err, cap):
        self.calls.append((list(offsets[:count + 1]), mode))
        out = bytes(self.verdicts[:count])
        self.verdicts = self.verdicts[count:]
        ctypes.memmove(res, out, count)
        return 0


class StubBackend:
    def __init__(self, verdicts):
        self.lib = StubLib(verdicts)
        self.handle = None


def test_mode_values_match_the_header():
    text = open(os.path.join(ROOT, "include", "dynaverse_pairing.h")).read()
    defs = dict(re.findall(r"#define DVGPU_PAIRING_(\w+) (\d+)", text))
    assert int(defs["PRODUCT"]) == VerificationMode.PRODUCT
    assert int(defs["REFERENCE"]) == VerificationMode.REFERENCE


def test_encoding_puts_the_left_pair_first():
    eqs = [((P, Q), [(None, Q)]), ((None, None), []), ((P, None), [(P, Q), (P, Q), (None, Q)])]
    g1, g2, off = encode_equations(eqs)
    assert off == [0, 2, 3, 7]
    assert g1[:64] == encode_g1(P) and g1[64:128] == bytes(64) and g2[:128] == encode_g2(Q)
    assert len(g1) == 64 * 7 and len(g2) == 128 * 7
    with pytest.raises(ValueError):
        encode_equations([((P, Q), [(P, Q)] * MAX_PAIRS)])


@pytest.mark.parametrize("mode", list(VerificationMode))
def test_trace_counts_follow_the_reference(mode):
    """verification.rs:38-45: Reference mode counts one Miller loop and one
    final exponentiation per pair, Product mode one of each per equation."""
    b = StubBackend([1, 0, 1, 1, 0, 1])
    trace = VerificationTrace()
    checks = PairingChecks(b, mode, trace)
    started = checks.begin(VerificationPhase.ANCHOR_RELATIONS)
    assert checks.equal((P, Q), [(P, Q)]) is True
    assert checks.equal((P, Q), [(P, Q), (P, Q)]) is False
    checks.end(started)
    checks.begin(VerificationPhase.DELTA_OPENINGS)
    assert checks.defer((P, Q), [(P, Q)] * 4) == 0
    assert checks.defer((P, Q), []) == 1
    checks.begin(VerificationPhase.PUBLIC_INPUTS)
    assert checks.defer((P, Q), [(P, Q)]) == 2
    assert checks.run() == [True, True, False]
    assert checks.run() == []
    m = trace.measurements
    per = (lambda t: t) if mode == VerificationMode.REFERENCE else (lambda t: 1)
    a = m[VerificationPhase.ANCHOR_RELATIONS].counts
    assert (a.equations, a.pairing_terms, a.miller_loops, a.final_exponentiations) == \
        (2, 5, per(2) + per(3), per(2) + per(3))
    d = m[VerificationPhase.DELTA_OPENINGS].counts
    assert (d.equations, d.pairing_terms, d.miller_loops) == (2, 6, per(5) + per(1))
    p = m[VerificationPhase.PUBLIC_INPUTS].counts
    assert (p.equations, p.pairing_terms, p.miller_loops) == (1, 2, per(2))
    assert m[VerificationPhase.ANCHOR_RELATIONS].elapsed > 0
    assert trace.totals().equations == 5
    # three device calls: two single equations, one batch of three
    assert [c[0] for c in b.lib.calls] == [[0, 2], [0, 3], [0, 5, 6, 8]]
    assert {c[1] for c in b.lib.calls} == {int(mode)}


def test_equal_prepared_is_a_two_pair_product():
    """verification.rs:55-68: one equation, two pairs, one Miller loop, one
    final exponentiation; Product mode only."""
    b = StubBackend([1])
    trace = VerificationTrace()
    checks = PairingChecks(b, VerificationMode.PRODUCT, trace)
    assert checks.equal_prepared(P, P, Q, Q)
    c = trace.measurements[VerificationPhase.STRUCTURE].counts
    assert (c.equations, c.pairing_terms, c.miller_loops, c.final_exponentiations) == (1, 2, 1, 1)
    assert b.lib.calls == [([0, 2], 0)]
    with pytest.raises(ValueError):
        PairingChecks(b, VerificationMode.REFERENCE).equal_prepared(P, P, Q, Q)


def test_a_disabled_trace_counts_nothing():
    b = StubBackend([1, 1])
    trace = VerificationTrace(enabled=False)
    checks = PairingChecks(b, VerificationMode.PRODUCT, trace)
    assert checks.begin(VerificationPhase.STRUCTURE) is None
    checks.equal((P, Q), [(P, Q)])
    checks.defer((P, Q), [])
    checks.run()
    assert trace.measurements == {}
