"""Host logic of the serving form (no GPU): range allocation, update merging,
seeded request streams and session derivation."""
import numpy as np

from dvgpu.codec import FR_MODULUS as R, root_of_unity_2_20
from dvgpu.serving import (RangeAllocator, SessionSpec, merge_update, percentile, request_stream,
                           session_bytes, session_pool, session_scalars)


def test_range_allocator_first_fit_and_merge():
    a = RangeAllocator(100)
    x, y, z = a.alloc(30), a.alloc(30), a.alloc(30)
    assert (x, y, z) == (0, 30, 60) and a.alloc(20) is None
    a.release(y, 30)
    assert a.alloc(40) is None and a.alloc(30) == 30
    a.release(x, 30)
    a.release(30, 30)
    a.release(z, 30)
    assert a.free == [(0, 100)] and a.free_units == 100


def test_merge_update_sums_repeats():
    idx, raw, vals = merge_update([5, 2, 5, 9, 2], [1, 2, 3, R - 1, R - 2])
    assert idx.tolist() == [2, 5, 9]
    assert vals == [0, 4, R - 1]
    assert raw == b"".join(v.to_bytes(32, "little") for v in vals)


def test_session_scalars_relations():
    spec = SessionSpec(0, 16, 123)
    tau, t, z, coeffs = spec.derive()
    w16 = pow(root_of_unity_2_20(), 1 << 16, R)
    assert pow(w16, 16, R) == 1 and z == pow(w16, t, R)
    a, u, l = session_scalars(tau, z, 16)
    for i in range(16):
        ai = int.from_bytes(a[i].tobytes(), "little")
        ui = int.from_bytes(u[i].tobytes(), "little")
        li = int.from_bytes(l[i].tobytes(), "little")
        assert ai == pow(tau, i, R) and li == pow(z, i, R)
        assert ui * (tau - z) % R == (ai - li) % R        # U_i = (tau^i - z^i)/(tau - z)
    # Opening relation for f: sum f_i U_i = (f(tau) - f(z))/(tau - z).
    f = [int.from_bytes(coeffs[i].tobytes(), "little") for i in range(16)]
    ft = sum(c * pow(tau, i, R) for i, c in enumerate(f)) % R
    fz = sum(c * pow(z, i, R) for i, c in enumerate(f)) % R
    q = sum(c * int.from_bytes(u[i].tobytes(), "little") for i, c in enumerate(f)) % R
    assert q * (tau - z) % R == (ft - fz) % R


def test_request_stream_deterministic_and_mixed():
    specs = session_pool(4)
    a = request_stream(specs, 400, 1000.0)
    b = request_stream(specs, 400, 1000.0)
    assert [(x[0], x[1], list(x[2]), list(x[3])) for x in a] == [(x[0], x[1], list(x[2]), list(x[3])) for x in b]
    sizes = {s.sid: s.n for s in specs}
    counts = {n: 0 for n in (256, 1024, 4096, 65536)}
    for t, sid, idx, ds in a:
        counts[sizes[sid]] += 1
        assert 1 <= len(idx) <= 64 and all(0 <= i < sizes[sid] for i in idx)
        assert all(0 <= d < R for d in ds)
    assert counts[256] > counts[1024] > counts[4096] > counts[65536] > 0
    gaps = np.diff([0.0] + [x[0] for x in a])
    assert abs(gaps.mean() - 1e-3) < 2.5e-4                # Poisson at 1000 QPS


def test_session_bytes_and_percentile():
    assert session_bytes(256) == 64 * 32 * 514 + 32 * 769
    assert percentile([1, 2, 3, 4, 5], 50) == 3 and percentile([1, 2, 3, 4, 5], 100) == 5


def test_coset_point_constant_is_off_subgroup():
    from dvgpu.benchmarks import COSET_POINT
    from oracle import bn254 as o
    from tests.g2corpus import twist_cofactor_point
    T = twist_cofactor_point()
    assert o.g2_to_bytes(T) == COSET_POINT
    assert o.on_curve(T, o.FQ2) and o.mul(T, o.R, o.FQ2) is not None


# ---- the request loop with stand-in lanes (no device) --------------------------

import threading
import time
from types import SimpleNamespace

from dvgpu.serving import Server, request_cost


class FakeStore:
    def __init__(self, sizes):
        self.sessions = {sid: SimpleNamespace(n=n, busy=False, divisor=sid.to_bytes(4, "little"))
                         for sid, n in enumerate(sizes)}


class FakeLane:
    """Applies updates to a per-session log; the 'proof' is the log length."""
    log = None
    lock = threading.Lock()
    active = 0

    def __init__(self, delay=0.002, poison=None):
        self.delay, self.poison = delay, poison

    def run(self, store, batch):
        sids = [r.sid for r in batch]
        assert len(set(sids)) == len(sids), "a session twice in one batch"
        if self.poison is not None and any(int(r.indices[0]) == self.poison for r in batch):
            raise RuntimeError("poisoned batch")
        with FakeLane.lock:
            FakeLane.active += 1
        time.sleep(self.delay)
        with FakeLane.lock:
            FakeLane.active -= 1
        out = []
        with FakeLane.lock:
            for r in batch:
                FakeLane.log.setdefault(r.sid, []).append(int(r.indices[0]))
                out.append(len(FakeLane.log[r.sid]).to_bytes(4, "little"))
        FakeLane.batches.append(len(batch))
        return out, 0.0

    def close(self):
        pass


class FakeVerifier:
    """Rejects proofs whose session id is in `bad`; records call sizes and
    checks that each proof comes with its own session's divisor."""

    def __init__(self, bad=(), delay=0.003):
        self.bad, self.delay, self.calls = set(bad), delay, []

    def run(self, proofs, divisors):
        assert FakeLane.active == 0, "verification overlapped a lane batch"
        time.sleep(self.delay)
        assert FakeLane.active == 0, "a lane batch started during verification"
        self.calls.append(len(proofs))
        return [int.from_bytes(d, "little") not in self.bad for d in divisors]

    def close(self):
        pass


def _serve(store, reqs, **kw):
    FakeLane.log, FakeLane.batches = {}, []
    lane_kw = kw.pop("lane", {})
    kw.setdefault("verify", "verifier_factory" in kw)
    srv = Server(store, lane_factory=lambda: FakeLane(**lane_kw), **kw)
    try:
        handles = [srv.submit(sid, np.array([seq], dtype=np.uint32), b"") for sid, seq in reqs]
        results = []
        for h in handles:
            try:
                results.append(h.future.result(timeout=30))
            except RuntimeError as e:
                results.append(e)
        return srv, handles, results
    finally:
        srv.close()


def test_server_keeps_session_order_and_shares_batches():
    store = FakeStore([256] * 40)
    rng = np.random.default_rng(3)
    reqs = [(int(rng.integers(0, 40)), i) for i in range(600)]
    srv, handles, results = _serve(store, reqs, lanes=2, max_batch=64)
    for sid in range(40):
        mine = [seq for s, seq in reqs if s == sid]
        assert FakeLane.log.get(sid, []) == mine               # arrival order per session
    # Each proof is its request's position in its session's sequence.
    pos = {}
    for (sid, _), res in zip(reqs, results):
        pos[sid] = pos.get(sid, 0) + 1
        assert int.from_bytes(res, "little") == pos[sid]
    assert srv.batches < len(reqs) and max(FakeLane.batches) > 1   # continuous batching
    assert all(h.done >= h.started >= h.submitted for h in handles)


def test_server_batch_limit_and_scratch_budget():
    store = FakeStore([65536] * 30)
    reqs = [(i % 30, i) for i in range(300)]
    budget = 3 * request_cost(65536, 1)                      # three large requests per lane batch
    srv, _, _ = _serve(store, reqs, lanes=2, max_batch=64, scratch_budget=2 * budget)
    assert max(FakeLane.batches) <= 3
    srv2, _, _ = _serve(FakeStore([256] * 100), [(i, i) for i in range(100)], lanes=1, max_batch=8,
                        lane={"delay": 0.01})
    assert max(FakeLane.batches) <= 8


def test_server_delivers_batch_errors_and_keeps_serving():
    store = FakeStore([256] * 8)
    reqs = [(i % 8, i) for i in range(80)]
    srv, handles, results = _serve(store, reqs, lanes=2, max_batch=16, lane={"poison": 13})
    failed = {seq for (sid, seq), r in zip(reqs, results) if isinstance(r, RuntimeError)}
    assert 13 in failed and 1 <= len(failed) <= 8             # one batch failed, with all its requests
    assert all(not isinstance(r, RuntimeError) for (sid, seq), r in zip(reqs, results) if seq > 40)
    assert not any(s.busy for s in store.sessions.values())   # sessions released after the error


def test_update_record_layout_matches_struct():
    import ctypes
    from dvgpu import _lib
    from dvgpu.serving import UPDATE_DTYPE
    assert UPDATE_DTYPE.itemsize == ctypes.sizeof(_lib.DvgpuUpdate)
    for name, _ in _lib.DvgpuUpdate._fields_:
        assert UPDATE_DTYPE.fields[name][1] == getattr(_lib.DvgpuUpdate, name).offset, name


def test_server_verifies_every_proof_before_resolving():
    store = FakeStore([256] * 12)
    ver = FakeVerifier(bad={5})
    reqs = [(i % 12, i) for i in range(240)]
    srv, handles, results = _serve(store, reqs, lanes=2, max_batch=8, verifier_factory=lambda: ver, verify_min=40)
    from dvgpu.serving import VerificationError
    for (sid, _), h, r in zip(reqs, handles, results):
        if sid == 5:
            assert isinstance(r, VerificationError)
        else:
            assert isinstance(r, bytes)
        assert h.verified >= h.done > 0 and h.finished == h.verified
    assert srv.verified == 220 and srv.rejected == 20
    assert sum(ver.calls) == 240 and len(ver.calls) == srv.verify_calls
    assert max(ver.calls) > 8            # one call verifies several finished batches


def test_server_close_fails_queued_requests():
    """Requests still queued when the server closes get an exception instead
    of staying pending for ever (one lane, one request per batch)."""
    FakeLane.log, FakeLane.batches = {}, []
    store = FakeStore([256])
    srv = Server(store, lanes=1, max_batch=1, verify=False, lane_factory=lambda: FakeLane(delay=0.05))
    handles = [srv.submit(0, np.array([i], dtype=np.uint32), b"") for i in range(4)]
    time.sleep(0.01)
    srv.close()
    outcomes = []
    for h in handles:
        try:
            outcomes.append(h.future.result(timeout=5))
        except RuntimeError as e:
            outcomes.append(str(e))
    assert isinstance(outcomes[0], bytes)               # the batch on the lane finished
    assert outcomes[1:] == ["server closed"] * 3
