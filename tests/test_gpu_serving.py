"""Serving form on the MI355X: batched sparse updates of device-resident
sessions, checked bit-exact against the one-call-at-a-time path and in
closed form through the SRS trapdoor (C = [f(tau)] g1, W = [(f(tau) - y) /
(tau - z)] g1, y = f(z) = NTT(f)[t]: the exponent form of the KZG pairing
check), plus the C oracle's NTT of the coefficients.  The served stream's
results equal the same requests run alone.
"""
import time

import numpy as np
import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.codec import FR_MODULUS as R, encode_fr, root_of_unity_2_20
from dvgpu.serving import (PlainSession, Server, SessionSpec, SessionStore, merge_update, percentile,
                           request_stream, session_bytes, session_pool, update_batch, verify_served)
from oracle import bn254 as o
from oracle import cref

pytestmark = pytest.mark.gpu
ROOT = encode_fr(root_of_unity_2_20())


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


class Mirror:
    """Host ints of one session's coefficients, for closed-form checks."""

    def __init__(self, spec):
        self.tau, self.t, self.z, coeffs = spec.derive()
        self.n = spec.n
        self.f = [int.from_bytes(coeffs[i].tobytes(), "little") for i in range(spec.n)]

    def apply(self, idx, dvals):
        for i, d in zip(idx, dvals):
            self.f[int(i)] = (self.f[int(i)] + d) % R

    def expect(self):
        ft = fz = 0
        for c in reversed(self.f):
            ft = (ft * self.tau + c) % R
            fz = (fz * self.z + c) % R
        q = (ft - fz) * pow((self.tau - self.z) % R, R - 2, R) % R
        g = lambda e: o.g1_to_bytes(o.g1_mul(e) if e else None)
        return g(ft) + g(q) + fz.to_bytes(32, "little")

    def coeff_bytes(self):
        return b"".join(c.to_bytes(32, "little") for c in self.f)


def _check_session(store, s, mirror, proof):
    assert proof == mirror.expect()
    assert store.coefficients(s) == mirror.coeff_bytes()
    ev = store.evaluations(s)
    assert ev == cref.ntt(mirror.coeff_bytes(), mirror.n, False, ROOT, 8)
    assert ev[32 * mirror.t:32 * mirror.t + 32] == proof[128:160]      # y = NTT(f)[t]


def test_anchor_and_batched_updates_closed_form(backend):
    store = SessionStore(backend, 1 << 30)
    specs = [SessionSpec(i, n, 900 + i) for i, n in enumerate((256, 1024, 4096, 65536, 256, 2))]
    sessions, mirrors = [], []
    for sp in specs:
        print(f"case: admit n={sp.n}", flush=True)
        sessions.append(store.admit(sp))
        mirrors.append(Mirror(sp))
    # Anchor state: an update with k = 0 returns the cached proof unchanged.
    keep = []
    ups = []
    for s in sessions:
        idx = np.zeros(0, dtype=np.uint32)
        keep.append(idx)
        ups.append(s.update_struct(idx, b""))
    print("case: k=0 batch", flush=True)
    proofs = update_batch(backend, store, ups)
    for s, m, p in zip(sessions, mirrors, proofs):
        _check_session(store, s, m, p)
    # Three rounds of batched updates with repeats (merged) over all sessions.
    rng = np.random.default_rng(7)
    for rnd in range(3):
        ups, keep = [], []
        for s, m in zip(sessions, mirrors):
            k = int(rng.integers(1, min(64, s.n) + 1))
            idx_raw = rng.integers(0, s.n, size=k).tolist() + [0]
            d_raw = [int(x) % R for x in rng.integers(0, 2**62, size=k + 1)]
            idx, raw, dvals = merge_update(idx_raw, d_raw)
            keep.append((idx, raw))
            ups.append(s.update_struct(idx, raw))
            m.apply(idx, dvals)
        print(f"case: update round {rnd}", flush=True)
        proofs, ms = update_batch(backend, store, ups, timed=True)
        assert ms > 0
        for s, m, p in zip(sessions, mirrors, proofs):
            _check_session(store, s, m, p)
    store.close()


def test_batch_rejections_leave_state(backend):
    store = SessionStore(backend, 16 << 20)
    s = store.admit(SessionSpec(0, 256, 77))
    m = Mirror(SessionSpec(0, 256, 77))
    dup = np.array([3, 3], dtype=np.uint32)
    raw = encode_fr(5) * 2
    with pytest.raises(BackendError):                   # repeated index inside one update
        update_batch(backend, store, [s.update_struct(dup, raw)])
    one = np.array([3], dtype=np.uint32)
    with pytest.raises(BackendError):                   # the same session twice in one batch
        update_batch(backend, store, [s.update_struct(one, raw[:32]), s.update_struct(one, raw[:32])])
    bad = np.array([256], dtype=np.uint32)
    with pytest.raises(BackendError):                   # index >= n
        update_batch(backend, store, [s.update_struct(bad, raw[:32])])
    with pytest.raises(BackendError):                   # non-canonical delta
        update_batch(backend, store, [s.update_struct(one, (R).to_bytes(32, "little"))])
    # Nothing above touched the session.
    p = update_batch(backend, store, [s.update_struct(np.zeros(0, dtype=np.uint32), b"")])[0]
    _check_session(store, s, m, p)
    store.close()


def test_hbm_budget_admission(backend):
    budget = 6 * session_bytes(1024) + 4096
    store = SessionStore(backend, budget)
    admitted = []
    for i in range(8):
        s = store.admit(SessionSpec(i, 1024, 500 + i))
        if s is None:
            break
        admitted.append(s)
    assert 2 <= len(admitted) < 8                       # the budget, not the loop, stops admission
    assert store.admit(SessionSpec(99, 1024, 599)) is None
    store.release(admitted[0])
    again = store.admit(SessionSpec(100, 1024, 600))
    assert again is not None
    m = Mirror(SessionSpec(100, 1024, 600))
    p = update_batch(backend, store, [again.update_struct(np.zeros(0, dtype=np.uint32), b"")])[0]
    _check_session(store, again, m, p)
    store.close()


def test_served_stream_equals_alone_and_plain(backend):
    specs = session_pool(2, seed=11)                    # 8 sessions over the size mix
    stream = request_stream(specs, 120, 0.0, seed=11)   # all requests at once
    # Served, continuous batching over two lanes.
    store = SessionStore(backend, 2 << 30)
    for sp in specs:
        store.admit(sp)
    print("case: served stream", flush=True)
    srv = Server(store, lanes=2, max_batch=64)
    reqs = []
    for t, sid, idx, ds in stream:
        mi, raw, _ = merge_update(idx, ds)
        reqs.append(srv.submit(sid, mi, raw))
    served = [r.future.result(timeout=120) for r in reqs]
    assert srv.batches < len(reqs)                      # requests shared batches
    srv.close()
    # every served proof passes its pairing check (the caller's owner, server closed)
    print("case: verify served proofs", flush=True)
    ok, _ = verify_served(backend, store, reqs)
    assert ok == [True] * len(reqs)
    ev_served = {sp.sid: store.evaluations(store.sessions[sp.sid]) for sp in specs}
    store.close()
    # Alone: a fresh store, every request in a batch of its own, in order.
    store2 = SessionStore(backend, 2 << 30)
    for sp in specs:
        store2.admit(sp)
    print("case: alone", flush=True)
    alone = []
    for t, sid, idx, ds in stream:
        mi, raw, _ = merge_update(idx, ds)
        alone.append(update_batch(backend, store2, [store2.sessions[sid].update_struct(mi, raw)])[0])
    assert served == alone
    for sp in specs:
        assert store2.evaluations(store2.sessions[sp.sid]) == ev_served[sp.sid]
    store2.close()
    # Plain one-call-at-a-time path through the reference operator API.
    print("case: plain path", flush=True)
    plain = {sp.sid: PlainSession(backend, sp) for sp in specs}
    got = []
    for t, sid, idx, ds in stream:
        mi, raw, dv = merge_update(idx, ds)
        got.append(plain[sid].apply(backend, mi, raw, dv))
    assert got == served
    for sp in specs:
        assert plain[sp.sid].evals == ev_served[sp.sid]
    # Closed form on the final state of every session.
    for sp in specs:
        m = Mirror(sp)
        for t, sid, idx, ds in stream:
            if sid == sp.sid:
                mi, raw, dv = merge_update(idx, ds)
                m.apply(mi, dv)
        last = [p for (t, sid, _, _), p in zip(stream, served) if sid == sp.sid]
        if last:
            assert last[-1] == m.expect()
