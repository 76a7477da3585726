"""Per-equation pairing checks on the device: the host mirror of
PairingChecks (crates/protocol/src/verification.rs:12-144) over
dvgpu_bn254_pairing_equations (include/dynaverse_pairing.h).

    checks = PairingChecks(backend, VerificationMode.PRODUCT, trace)
    started = checks.begin(VerificationPhase.ANCHOR_RELATIONS)
    ok = checks.equal((A, B), [(C, D)]) and checks.equal((E, F), [(G, H), (I, J)])
    checks.end(started)

An equation is e(L.P, L.Q) == prod_j e(R_j.P, R_j.Q) with affine points
(int tuples, None = identity, as everywhere in this package).  `equal`
evaluates one equation at once, so a caller's short-circuit order keeps its
meaning and the trace counts what the reference counts
(verification.rs:37-45).  `defer` queues an equation and `run` evaluates every
queued equation in one device call; each is still decided on its own (the
reference never merges distinct equations, verification.rs:1, and neither
does this).  The modes decide identically; Reference mode computes every
pairing separately, as the reference's cross-check path does.
"""
import ctypes
import enum
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from .backend import HipBackend, _check, _err
from .codec import encode_g1, encode_g2

MAX_PAIRS = 16  # pairs per equation, the device bound


class VerificationMode(enum.IntEnum):
    """types.rs:193-196; values are DVGPU_PAIRING_PRODUCT / _REFERENCE."""
    PRODUCT = 0
    REFERENCE = 1


class VerificationPhase(enum.Enum):
    """trace.rs:167-178."""
    STRUCTURE = "structure"
    ANCHOR_DYNAMO = "anchor_dynamo"
    ANCHOR_CONSISTENCY = "anchor_consistency"
    ANCHOR_RELATIONS = "anchor_relations"
    DELTA_DYNAMO = "delta_dynamo"
    DELTA_CONSISTENCY = "delta_consistency"
    DELTA_DIVISORS = "delta_divisors"
    DELTA_MULTIPLICATION = "delta_multiplication"
    DELTA_OPENINGS = "delta_openings"
    PUBLIC_INPUTS = "public_inputs"


@dataclass
class PairingCounts:
    """trace.rs:208-222."""
    equations: int = 0
    pairing_terms: int = 0
    miller_loops: int = 0
    final_exponentiations: int = 0

    def add(self, other: "PairingCounts") -> None:
        self.equations += other.equations
        self.pairing_terms += other.pairing_terms
        self.miller_loops += other.miller_loops
        self.final_exponentiations += other.final_exponentiations


@dataclass
class VerificationMeasurement:
    """trace.rs:224-236, with the device time of this phase's calls in place
    of the host Miller / final-exponentiation split."""
    elapsed: float = 0.0      # seconds between begin and end
    device_ms: float = 0.0    # event-timed kernel time of the phase's calls
    calls: int = 0            # device calls
    counts: PairingCounts = field(default_factory=PairingCounts)


@dataclass
class VerificationTrace:
    enabled: bool = True
    measurements: Dict[VerificationPhase, VerificationMeasurement] = field(default_factory=dict)

    def start(self) -> Optional[float]:
        return time.perf_counter() if self.enabled else None

    def measurement_mut(self, phase: VerificationPhase) -> VerificationMeasurement:
        return self.measurements.setdefault(phase, VerificationMeasurement())

    def totals(self) -> PairingCounts:
        out = PairingCounts()
        for m in self.measurements.values():
            out.add(m.counts)
        return out


Pair = Tuple[object, object]
Equation = Tuple[Pair, Sequence[Pair]]


def encode_equations(equations: Sequence[Equation]) -> Tuple[bytes, bytes, List[int]]:
    """g1 bytes, g2 bytes and the count + 1 offsets of a batch, the left pair
    of every equation first."""
    g1, g2, offsets = [], [], [0]
    for left, right in equations:
        pairs = [left, *right]
        if len(pairs) > MAX_PAIRS:
            raise ValueError("an equation has more than %d pairs" % MAX_PAIRS)
        for p, q in pairs:
            g1.append(encode_g1(p))
            g2.append(encode_g2(q))
        offsets.append(offsets[-1] + len(pairs))
    return b"".join(g1), b"".join(g2), offsets


def pairing_equations_encoded(backend: HipBackend, g1: bytes, g2: bytes, offsets: Sequence[int],
                              mode: VerificationMode = VerificationMode.PRODUCT):
    """Encoded form: (results bytes, kernel ms) for len(offsets) - 1 equations."""
    count = len(offsets) - 1
    if count <= 0:
        return b"", 0.0
    off = (ctypes.c_uint32 * len(offsets))(*offsets)
    res = ctypes.create_string_buffer(count)
    ms = ctypes.c_float(0)
    err = _err()
    _check(backend.lib.dvgpu_bn254_pairing_equations(backend.handle, ctypes.c_char_p(g1 or b"\0"),
                                                     ctypes.c_char_p(g2 or b"\0"), off, count, int(mode), res,
                                                     ctypes.byref(ms), err, 256), err)
    return res.raw[:count], ms.value


def pairing_equations(backend: HipBackend, equations: Sequence[Equation],
                      mode: VerificationMode = VerificationMode.PRODUCT) -> List[bool]:
    """Whether each equation holds, all in one device call."""
    if not equations:
        return []
    g1, g2, offsets = encode_equations(equations)
    res, _ = pairing_equations_encoded(backend, g1, g2, offsets, mode)
    return [b == 1 for b in res]


class PairingChecks:
    """verification.rs:12-144: one checker per verification, bound to one
    trace.  The reference caches G2 preparations per checker; the device
    prepares every G2 point inside its Miller loop, so there is no cache."""

    def __init__(self, backend: HipBackend, mode: VerificationMode = VerificationMode.PRODUCT,
                 trace: Optional[VerificationTrace] = None):
        self.backend = backend
        self.mode = VerificationMode(mode)
        self.trace = trace if trace is not None else VerificationTrace(enabled=False)
        self.phase = VerificationPhase.STRUCTURE
        self._queued: List[Tuple[Equation, VerificationPhase]] = []

    def begin(self, phase: VerificationPhase) -> Optional[float]:
        """verification.rs:26-29."""
        self.phase = phase
        return self.trace.start()

    def end(self, started: Optional[float]) -> None:
        """verification.rs:31-35."""
        if started is not None:
            self.trace.measurement_mut(self.phase).elapsed += time.perf_counter() - started

    def _count(self, phase: VerificationPhase, terms: int, ms: float = 0.0, calls: int = 0) -> None:
        if not self.trace.enabled:
            return
        loops = terms if self.mode == VerificationMode.REFERENCE else 1
        m = self.trace.measurement_mut(phase)
        m.counts.add(PairingCounts(1, terms, loops, loops))
        m.device_ms += ms
        m.calls += calls

    def _evaluate(self, equations: Sequence[Equation], mode: VerificationMode) -> Tuple[List[bool], float]:
        g1, g2, offsets = encode_equations(equations)
        res, ms = pairing_equations_encoded(self.backend, g1, g2, offsets, mode)
        return [b == 1 for b in res], ms

    def equal(self, left: Pair, right: Sequence[Pair]) -> bool:
        """e(left) == prod e(right_j), evaluated now (verification.rs:37-53)."""
        (ok,), ms = self._evaluate([(left, right)], self.mode)
        self._count(self.phase, 1 + len(right), ms, 1)
        return ok

    def equal_prepared(self, left, right, generator, divisor) -> bool:
        """e(left, generator) == e(right, divisor) as one two-pair product
        (verification.rs:55-77; the reference allows Product mode only)."""
        if self.mode != VerificationMode.PRODUCT:
            raise ValueError("equal_prepared is a Product-mode check")
        (ok,), ms = self._evaluate([((left, generator), [(right, divisor)])], VerificationMode.PRODUCT)
        if self.trace.enabled:
            m = self.trace.measurement_mut(self.phase)
            m.counts.add(PairingCounts(1, 2, 1, 1))
            m.device_ms += ms
            m.calls += 1
        return ok

    def defer(self, left: Pair, right: Sequence[Pair]) -> int:
        """Queue e(left) == prod e(right_j) under the current phase; returns
        its index in the next run()."""
        if 1 + len(right) > MAX_PAIRS:
            raise ValueError("an equation has more than %d pairs" % MAX_PAIRS)
        self._queued.append(((left, list(right)), self.phase))
        return len(self._queued) - 1

    def run(self) -> List[bool]:
        """Every queued equation in one device call, each decided on its own;
        the trace counts each as equal() would.  The queue is emptied, also
        when the call fails (the error, e.g. a point off its curve, is raised
        and nothing is counted)."""
        queued, self._queued = self._queued, []
        if not queued:
            return []
        out, ms = self._evaluate([e for e, _ in queued], self.mode)
        for k, ((_, right), phase) in enumerate(queued):
            self._count(phase, 1 + len(right), ms if k == 0 else 0.0, 1 if k == 0 else 0)
        return out
