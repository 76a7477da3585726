"""Serving form: a request loop over device-resident session states with
continuous batching and HBM-budget admission.

A session (one Dynaverse state of size n) keeps everything an update needs in
HBM, inside two arenas carved out of a fixed memory budget:

* its own seeded SRS basis P_i = [tau^i] g1 and opening basis
  U_i = [(tau^i - z^i) / (tau - z)] g1, frozen, as window tables: row i holds
  [2^(c w)] P_i for w < ceil(255 / c) (built on the device by the fixed-base
  kernel), so an update's MSM needs no doublings between windows;
* the cached commitment C = [f(tau)] g1 and opening W = [q(tau)] g1 with
  q = (f - f(z)) / (X - z), as XYZZ points updated in place;
* coefficients f, evaluations NTT(f), the frozen linear term l_i = z^i and the
  cached value y = f(z).

An update request changes k coefficients; its proof is (C', W', y'), which
satisfies the KZG opening relation f'(tau) - y' = q'(tau) (tau - z), checked
as the reference verifies an update (crates/protocol/src/dynaverse.rs:
447-451): e(C' - [y'] g1, g2) == e(W', D) with the session's frozen G2 term
D = [tau - z] g2, on the device (dvgpu_bn254_kzg_verify_batch).  One
``dvgpu_ext_update_batch`` call serves a whole batch of requests (two sparse
window-table MSMs per request, scatter-add, linear term, a fused
multi-length NTT).  The ``Server`` keeps ``lanes`` such calls in flight, each
on its own owner and stream; a request that arrives while others run joins
the next batch that a free lane takes (continuous batching).  Requests of one
session are applied in arrival order, one at a time.

Verification.  Every served proof is pairing-checked on the device by
``verify_served``: one dvgpu_bn254_kzg_verify_batch call on the caller's
owner over many proofs, after the server has closed.  The Server also has an
in-loop verification stage (a request then resolves only after its proof
passed; verification calls run while no lane batch is on the device, in
chunks of at least `verify_min` proofs), exercised on the host with a
stand-in verifier.  Its device verifier, a separate owner, is switched off:
its first pairing dispatch (a kernel with ~23 KB of scratch per lane) with
four owners alive in the process was refused by the runtime
(HSA_STATUS_ERROR_OUT_OF_RESOURCES) in two runs, while every pairing dispatch
made with the caller's owner alone alive has run.

``PlainSession`` is the one-call-at-a-time baseline through the reference's
operator API (``msm_batch`` over host-gathered bases, ``ntt`` on host data),
with the same state held on the host.  Both paths produce identical bytes.
"""
import collections
import ctypes
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .backend import BackendError, DeviceArray, HipBackend, _check, _err
from .codec import FR_MODULUS, encode_fr, root_of_unity_2_20
from .rng import ChaCha20Stream

R = FR_MODULUS
PROOF_BYTES = 160   # C' (64) || W' (64) || y' (32)
_ROOT = encode_fr(root_of_unity_2_20())


def _le_rows(vals: Sequence[int]) -> np.ndarray:
    return np.frombuffer(b"".join(int(v).to_bytes(32, "little") for v in vals), dtype=np.uint8).reshape(-1, 32)


@dataclass
class SessionSpec:
    """Seeded definition of one session: SRS trapdoor tau, opening point
    z = w_n^t (a domain point, so y = NTT(f)[t]) and initial coefficients."""
    sid: int
    n: int
    seed: int

    def derive(self):
        rng = ChaCha20Stream(self.seed)
        tau = rng.nonzero_fr()
        t = rng.below(self.n)
        w_n = pow(root_of_unity_2_20(), (1 << 20) // self.n, R)
        z = pow(w_n, t, R)
        coeffs = rng.uniform_fr_array(self.n)
        return tau, t, z, coeffs


def session_scalars(tau: int, z: int, n: int):
    """(tau^i, (tau^i - z^i)/(tau - z), z^i) for i < n as (n, 32) arrays."""
    inv = pow((tau - z) % R, R - 2, R)
    tp, zp = 1, 1
    a, b, c = [], [], []
    for _ in range(n):
        a.append(tp)
        b.append((tp - zp) * inv % R)
        c.append(zp)
        tp = tp * tau % R
        zp = zp * z % R
    return _le_rows(a), _le_rows(b), _le_rows(c)


class RangeAllocator:
    """First-fit allocator of [0, capacity) (unit-agnostic)."""

    def __init__(self, capacity: int):
        self.capacity = capacity
        self.free: List[Tuple[int, int]] = [(0, capacity)]     # sorted (start, size)

    def alloc(self, size: int) -> Optional[int]:
        for i, (s, z) in enumerate(self.free):
            if z >= size:
                if z == size:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + size, z - size)
                return s
        return None

    def release(self, start: int, size: int) -> None:
        self.free.append((start, size))
        self.free.sort()
        merged = []
        for s, z in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + z)
            else:
                merged.append((s, z))
        self.free = merged

    @property
    def free_units(self) -> int:
        return sum(z for _, z in self.free)


TABLE_BITS = 8                               # window width of the resident tables
TABLE_ROW = -(-255 // TABLE_BITS)            # points per table row


@dataclass
class Session:
    spec: SessionSpec
    log_n: int
    tau: int
    t: int
    z: int
    row: int              # points arena rows: [P table (n)] [U table (n)] [C] [W]
    fr_base: int          # fr: [f (n)] [evals (n)] [l (n)] [y]
    busy: bool = False
    record: tuple = ()
    divisor: bytes = b""  # [tau - z] g2, the verifier's frozen G2 term (128 B)

    @property
    def n(self):
        return self.spec.n

    def update_struct(self, idx: np.ndarray, deltas: bytes) -> _lib.DvgpuUpdate:
        n = self.n
        u = _lib.DvgpuUpdate()
        u.srs_row, u.opening_row = self.row, self.row + n
        u.commitment_row, u.opening_cache_row = self.row + 2 * n, self.row + 2 * n + 1
        u.coeff_offset, u.eval_offset = self.fr_base, self.fr_base + n
        u.linear_offset, u.value_slot = self.fr_base + 2 * n, self.fr_base + 3 * n
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        u.log_n, u.k = self.log_n, len(idx)
        u.indices = idx.ctypes.data if len(idx) else None
        u.deltas = ctypes.cast(ctypes.c_char_p(deltas), ctypes.c_void_p).value if len(idx) else None
        u._keep = (idx, deltas)       # the host arrays live as long as the struct
        return u


def session_divisor(backend: HipBackend, tau: int, z: int) -> bytes:
    """[tau - z] g2 (canonical, 128 B) through the setup entry point."""
    return backend.fixed_base_encoded(encode_fr((tau - z) % R), 1, True)


class VerificationError(RuntimeError):
    """A served proof failed its pairing check."""


def session_rows(n: int) -> int:
    return 2 * n + 2


def session_fr(n: int) -> int:
    return 3 * n + 1


def session_bytes(n: int) -> int:
    return 64 * TABLE_ROW * session_rows(n) + 32 * session_fr(n)


def merge_update(indices: Sequence[int], deltas: Sequence[int]) -> Tuple[np.ndarray, bytes, List[int]]:
    """Repeated indices add their deltas: returns sorted distinct indices,
    their canonical summed deltas (bytes) and the summed deltas (ints)."""
    acc: Dict[int, int] = {}
    for i, d in zip(indices, deltas):
        acc[int(i)] = (acc.get(int(i), 0) + int(d)) % R
    keys = sorted(acc)
    vals = [acc[k] for k in keys]
    return np.array(keys, dtype=np.uint32), b"".join(v.to_bytes(32, "little") for v in vals), vals


class SessionStore:
    """Session states in two HBM arenas under a fixed byte budget."""

    def __init__(self, backend: HipBackend, budget_bytes: int, point_share: float = 0.97):
        self.b = backend
        self.budget = int(budget_bytes)
        rows = int(self.budget * point_share) // (64 * TABLE_ROW)
        frs = (self.budget - rows * 64 * TABLE_ROW) // 32
        if rows * TABLE_ROW >= 1 << 32:
            raise ValueError("point arena limited to 2^32 points")
        self.points = backend.alloc(64 * TABLE_ROW * rows)
        self.fr = backend.alloc(32 * frs)
        self.point_count, self.fr_count = rows * TABLE_ROW, frs
        self.row_alloc, self.fr_alloc = RangeAllocator(rows), RangeAllocator(frs)
        self.sessions: Dict[int, Session] = {}

    def can_admit(self, n: int) -> bool:
        return any(z >= session_rows(n) for _, z in self.row_alloc.free) and \
            any(z >= session_fr(n) for _, z in self.fr_alloc.free)

    def admit(self, spec: SessionSpec) -> Optional[Session]:
        """Build the session's window tables on the device and prove its
        anchor state; None when the budget has no room."""
        n = spec.n
        if n & (n - 1) or not 1 <= n <= 1 << 20:
            raise ValueError("session size must be a power of two <= 2^20")
        row = self.row_alloc.alloc(session_rows(n))
        if row is None:
            return None
        fb = self.fr_alloc.alloc(session_fr(n))
        if fb is None:
            self.row_alloc.release(row, session_rows(n))
            return None
        tau, t, z, coeffs = spec.derive()
        s = Session(spec, n.bit_length() - 1, tau, t, z, row, fb)
        s.record = session_record(s)
        s.divisor = session_divisor(self.b, tau, z)
        a, u, l = session_scalars(tau, z, n)
        tmp = self.b.to_device(np.concatenate([a.reshape(-1), u.reshape(-1)]))
        rb = 64 * TABLE_ROW
        for k, base in ((0, row), (1, row + n)):
            err = _err()
            _check(self.b.lib.dvgpu_ext_window_table_device(self.b.handle, tmp.ptr + 32 * n * k, n, TABLE_BITS,
                                                            self.points.ptr + rb * base, err, 256), err)
        tmp.free()
        for r in (row + 2 * n, row + 2 * n + 1):
            self.points.upload_at(rb * r, np.zeros(128, dtype=np.uint8))
        self.fr.upload_at(32 * fb, np.zeros(32 * 2 * n, dtype=np.uint8))
        self.fr.upload_at(32 * (fb + 2 * n), l.reshape(-1))
        self.fr.upload_at(32 * (fb + 3 * n), np.zeros(32, dtype=np.uint8))
        # Anchor: the initial coefficients enter as one dense update.
        u_struct = s.update_struct(np.arange(n, dtype=np.uint32), coeffs.tobytes())
        update_batch(self.b, self, [u_struct])
        self.sessions[spec.sid] = s
        return s

    def release(self, s: Session) -> None:
        self.row_alloc.release(s.row, session_rows(s.n))
        self.fr_alloc.release(s.fr_base, session_fr(s.n))
        self.sessions.pop(s.spec.sid, None)

    def evaluations(self, s: Session) -> bytes:
        return self.fr.download_at(32 * (s.fr_base + s.n), 32 * s.n).tobytes()

    def coefficients(self, s: Session) -> bytes:
        return self.fr.download_at(32 * s.fr_base, 32 * s.n).tobytes()

    def close(self):
        self.points.free()
        self.fr.free()


def update_batch(backend: HipBackend, store: SessionStore, updates: Sequence[_lib.DvgpuUpdate],
                 timed: bool = False):
    """One native call for the batch: -> list of 160-byte proofs (and ms)."""
    arr = (_lib.DvgpuUpdate * len(updates))(*updates)
    return _update_call(backend, store, ctypes.cast(arr, ctypes.POINTER(_lib.DvgpuUpdate)), len(updates), timed)


def _update_call(backend, store, ptr, count, timed):
    out = ctypes.create_string_buffer(PROOF_BYTES * count)
    ms = ctypes.c_float(0)
    err = _err()
    _check(backend.lib.dvgpu_ext_update_batch(backend.handle, store.points.ptr, store.point_count, TABLE_BITS,
                                              store.fr.ptr, store.fr_count, ptr, count, _ROOT, out,
                                              ctypes.byref(ms) if timed else None, err, 256), err)
    raw = out.raw
    proofs = [raw[PROOF_BYTES * i:PROOF_BYTES * (i + 1)] for i in range(count)]
    return (proofs, ms.value) if timed else proofs


# DvgpuUpdate as a numpy record, for building a whole batch without one
# ctypes object per request (the layout is asserted against the struct).
UPDATE_DTYPE = np.dtype([("srs_row", "<u4"), ("opening_row", "<u4"), ("commitment_row", "<u4"),
                         ("opening_cache_row", "<u4"), ("coeff_offset", "<u8"), ("eval_offset", "<u8"),
                         ("linear_offset", "<u8"), ("value_slot", "<u8"), ("log_n", "<u4"), ("k", "<u4"),
                         ("indices", "<u8"), ("deltas", "<u8")])


def session_record(s: "Session") -> tuple:
    n = s.n
    return (s.row, s.row + n, s.row + 2 * n, s.row + 2 * n + 1, s.fr_base, s.fr_base + n, s.fr_base + 2 * n,
            s.fr_base + 3 * n, s.log_n, 0, 0, 0)


def update_batch_records(backend: HipBackend, store: SessionStore, batch: Sequence["Request"], timed: bool = False):
    """update_batch for queued requests, built as one numpy record array."""
    m = len(batch)
    rec = np.empty(m, dtype=UPDATE_DTYPE)
    sessions = store.sessions
    rec[:] = [sessions[r.sid].record for r in batch]
    keep = []
    ks = np.empty(m, dtype=np.uint32)
    ip = np.zeros(m, dtype=np.uint64)
    dp = np.zeros(m, dtype=np.uint64)
    for i, r in enumerate(batch):
        k = len(r.indices)
        ks[i] = k
        if k:
            idx = r.indices if r.indices.dtype == np.uint32 and r.indices.flags.c_contiguous else \
                np.ascontiguousarray(r.indices, dtype=np.uint32)
            keep.append(idx)
            ip[i] = idx.ctypes.data
            dp[i] = ctypes.cast(ctypes.c_char_p(r.deltas), ctypes.c_void_p).value
    rec["k"] = ks
    rec["indices"] = ip
    rec["deltas"] = dp
    return _update_call(backend, store, rec.ctypes.data_as(ctypes.POINTER(_lib.DvgpuUpdate)), m, timed)


# ---- request loop --------------------------------------------------------------

@dataclass
class Request:
    sid: int
    indices: np.ndarray
    deltas: bytes
    arrival: float = 0.0
    future: Future = field(default_factory=Future)
    submitted: float = 0.0
    started: float = 0.0
    done: float = 0.0        # proof computed
    verified: float = 0.0    # proof verified (0 when the server does not verify)

    @property
    def finished(self) -> float:
        return self.verified or self.done


def request_cost(n: int, k: int) -> int:
    """Scratch bytes a request adds to a batch (MSM entries at c = 4 and the
    NTT work row), for admission against the lanes' scratch budget."""
    return 2 * (k + 1) * 64 * 16 + 64 * n + 4096


class HipLane:
    """One worker lane of the Server: its own owner (and so its own stream),
    workspaces sized up front for the largest batch (one update per resident
    session at most), so serving allocates nothing: a growing device buffer
    would free the old one, and that synchronises the whole device, stalling
    every lane."""

    def __init__(self, store: SessionStore, max_batch: int, max_k: int):
        self.b = HipBackend()
        elements = sum(sess.n for sess in store.sessions.values()) or 1
        err = _err()
        _check(self.b.lib.dvgpu_ext_update_reserve(self.b.handle, max_batch, max_batch * max_k, elements,
                                                   TABLE_BITS, _ROOT, err, 256), err)

    def run(self, store: SessionStore, batch: Sequence["Request"]):
        """-> (one 160-byte proof per request, device ms)."""
        return update_batch_records(self.b, store, batch, timed=True)

    def close(self):
        self.b.close()


class VerifierLane:
    """The verification stage's device worker (its own owner and stream).
    Switched off: see the module docstring; use verify_served instead."""

    def __init__(self):
        raise NotImplementedError("the in-loop device verifier is switched off (the runtime refused its pairing "
                                  "dispatch with several owners alive); verify with verify_served after close()")

    def run(self, proofs: Sequence[bytes], divisors: Sequence[bytes]) -> List[bool]:
        from .pairing import verify_update_proofs
        return verify_update_proofs(self.b, proofs, divisors)

    def close(self):
        self.b.close()


class Server:
    """Continuous-batching request loop over a SessionStore.

    `lane_factory()` makes one worker lane (an object with run(store, batch)
    and close()); the default is a HipLane.  `verify` adds the verification
    stage (`verifier_factory()`, default VerifierLane: run(proofs, divisors)
    -> list of bools); each request then resolves after its proof passed, or
    with a VerificationError.  close() fails every request still queued with
    RuntimeError("server closed")."""

    def __init__(self, store: SessionStore, lanes: int = 2, max_batch: int = 256,
                 scratch_budget: int = 1 << 30, max_k: int = 64, lane_factory=None, verify: bool = False,
                 verifier_factory=None, max_verify: int = 1 << 16, verify_min: int = 1024):
        self.store = store
        self.lanes = lanes
        self.max_batch = max_batch
        self.lane_budget = scratch_budget // max(1, lanes)
        # Per-session FIFOs and the sessions ready to run (idle, with a
        # pending request), so forming a batch costs O(batch), not O(queue).
        self.pending: Dict[int, collections.deque] = collections.defaultdict(collections.deque)
        self.ready: collections.deque = collections.deque()
        self.cv = threading.Condition()
        self.stop_flag = False
        self.batches = 0
        self.batch_sizes: List[int] = []
        self.device_ms: List[float] = []
        factory = lane_factory or (lambda: HipLane(store, max_batch, max_k))
        self.workers = [factory() for _ in range(lanes)]
        self.work: List[collections.deque] = [collections.deque() for _ in range(lanes)]
        self.idle = list(range(lanes))
        # verification stage: finished batches wait here for one verify call;
        # while want_verify is set the scheduler starts no batch, and the
        # call runs once every lane is idle
        self.vqueue: collections.deque = collections.deque()
        self.vpending = 0
        self.vstop = False
        self.want_verify = False
        self.max_verify = max_verify
        self.verify_min = verify_min
        self.verify_calls = 0
        self.verify_sizes: List[int] = []
        self.verify_ms: List[float] = []
        self.verified = 0
        self.rejected = 0
        self.verifier = (verifier_factory or VerifierLane)() if verify else None
        self.threads = [threading.Thread(target=self._lane, args=(i,), daemon=True) for i in range(lanes)]
        self.sched = threading.Thread(target=self._schedule, daemon=True)
        self.vthread = threading.Thread(target=self._verify_loop, daemon=True) if verify else None
        for t in self.threads:
            t.start()
        self.sched.start()
        if self.vthread:
            self.vthread.start()

    def submit(self, sid: int, indices: np.ndarray, deltas: bytes, arrival: Optional[float] = None) -> Request:
        if sid not in self.store.sessions:
            raise KeyError("unknown session")
        now = time.perf_counter()
        req = Request(sid, indices, deltas, arrival if arrival is not None else now)
        req.submitted = now
        with self.cv:
            q = self.pending[sid]
            q.append(req)
            if len(q) == 1 and not self.store.sessions[sid].busy:
                self.ready.append(sid)
            self.cv.notify_all()
        return req

    def _take_batch(self) -> List[Request]:
        """One request from each of up to max_batch ready sessions, in the
        order they became ready, within the lane's scratch budget."""
        batch, cost = [], 0
        while self.ready and len(batch) < self.max_batch:
            sid = self.ready[0]
            req = self.pending[sid][0]
            c = request_cost(self.store.sessions[sid].n, len(req.indices))
            if batch and cost + c > self.lane_budget:
                break
            self.ready.popleft()
            self.pending[sid].popleft()
            self.store.sessions[sid].busy = True
            batch.append(req)
            cost += c
        return batch

    def _schedule(self):
        while True:
            with self.cv:
                while not self.stop_flag and (not self.idle or not self._runnable() or self.want_verify):
                    self.cv.wait()
                if self.stop_flag:
                    return
                batch = self._take_batch()
                if not batch:
                    continue
                lane = self.idle.pop()
                self.work[lane].append(batch)
                self.cv.notify_all()

    def _runnable(self) -> bool:
        return bool(self.ready)

    def _lane(self, i: int):
        worker = self.workers[i]
        while True:
            with self.cv:
                while not self.stop_flag and not self.work[i]:
                    self.cv.wait()
                if self.stop_flag and not self.work[i]:
                    return
                batch = self.work[i].popleft()
            t0 = time.perf_counter()
            err = None
            try:
                proofs, ms = worker.run(self.store, batch)
            except Exception as e:          # delivered to every request of the batch
                err, proofs, ms = e, None, 0.0
            t1 = time.perf_counter()
            with self.cv:
                for r in batch:
                    r.started, r.done = t0, t1
                    self.store.sessions[r.sid].busy = False
                    if self.pending[r.sid]:
                        self.ready.append(r.sid)
                self.batches += 1
                self.batch_sizes.append(len(batch))
                self.device_ms.append(ms)
                self.idle.append(i)
                self.cv.notify_all()
            if self.verifier is not None and err is None:
                with self.cv:
                    self.vqueue.append((batch, proofs))
                    self.vpending += len(batch)
                    self.cv.notify_all()
                continue
            for k, r in enumerate(batch):
                if err is not None:
                    r.future.set_exception(err)
                else:
                    r.future.set_result(proofs[k])

    def _verify_loop(self):
        """Takes every finished batch waiting (up to max_verify proofs) into
        one verification call, run while every lane is idle; resolves each
        request by its own result."""
        while True:
            with self.cv:
                while not (self.vqueue and (self.vpending >= self.verify_min or not self.ready or self.vstop)):
                    if self.vstop and not self.vqueue:
                        return
                    self.cv.wait()
                self.want_verify = True
                while len(self.idle) < self.lanes:
                    self.cv.wait()
                reqs, proofs = [], []
                while self.vqueue and (not reqs or len(reqs) + len(self.vqueue[0][0]) <= self.max_verify):
                    batch, pr = self.vqueue.popleft()
                    reqs.extend(batch)
                    proofs.extend(pr)
                self.vpending -= len(reqs)
            t0 = time.perf_counter()
            try:
                ok = self.verifier.run(proofs, [self.store.sessions[r.sid].divisor for r in reqs])
            except Exception as e:          # delivered to every request of the call
                ok = None
                failure = e
            t1 = time.perf_counter()
            with self.cv:
                self.want_verify = False
                self.cv.notify_all()
            if ok is None:
                for r in reqs:
                    r.future.set_exception(failure)
                continue
            self.verify_calls += 1
            self.verify_sizes.append(len(reqs))
            self.verify_ms.append((t1 - t0) * 1e3)
            for r, p, good in zip(reqs, proofs, ok):
                r.verified = t1
                if good:
                    self.verified += 1
                    r.future.set_result(p)
                else:
                    self.rejected += 1
                    r.future.set_exception(VerificationError(f"session {r.sid}: update proof failed its pairing check"))

    def close(self):
        """Stops taking batches; batches already on a lane finish (and are
        verified); every request still queued fails with "server closed"."""
        with self.cv:
            self.stop_flag = True
            self.cv.notify_all()
        self.sched.join()
        for t in self.threads:
            t.join()
        if self.vthread is not None:
            with self.cv:
                self.vstop = True
                self.cv.notify_all()
            self.vthread.join()
            self.verifier.close()
        for w in self.workers:
            w.close()
        with self.cv:
            left = [r for q in self.pending.values() for r in q]
            for q in self.pending.values():
                q.clear()
            self.ready.clear()
        for r in left:
            if not r.future.done():
                r.future.set_exception(RuntimeError("server closed"))


# ---- one-call-at-a-time baseline ----------------------------------------------

def _fixed_base_host(backend: HipBackend, scalars: np.ndarray) -> np.ndarray:
    """[s_i] g1 as canonical (n, 64) bytes through the setup entry point, in
    its 4096-point chunks (setup.rs:113-143)."""
    n = scalars.shape[0]
    out = [backend.fixed_base_encoded(scalars[c:c + 4096].tobytes(), min(4096, n - c), False)
           for c in range(0, n, 4096)]
    return np.frombuffer(b"".join(out), dtype=np.uint8).reshape(n, 64)


class PlainSession:
    """The same session state on the host, as the reference's plain path
    keeps it (canonical bases, coefficients, cached points and value)."""

    def __init__(self, backend: HipBackend, spec: SessionSpec):
        n = spec.n
        tau, t, z, coeffs = spec.derive()
        self.spec, self.n, self.t, self.z, self.tau = spec, n, t, z, tau
        self.divisor = session_divisor(backend, tau, z)
        a, u, l = session_scalars(tau, z, n)
        self.P = _fixed_base_host(backend, a)
        self.U = _fixed_base_host(backend, u)
        self.l = [int.from_bytes(l[i].tobytes(), "little") for i in range(n)]
        self.C = bytes(64)
        self.W = bytes(64)
        self.y = 0
        self.evals = b""
        idx = np.arange(n, dtype=np.uint32)
        dvals = [int.from_bytes(coeffs[i].tobytes(), "little") for i in range(n)]
        self.f = np.zeros_like(coeffs)
        self.apply(backend, idx, coeffs.tobytes(), dvals)

    def apply(self, backend: HipBackend, idx: np.ndarray, deltas: bytes, dvals: List[int]) -> bytes:
        """msm_batch over the gathered bases (+ the cached point with scalar
        1), the coefficient update on the host, ntt of the coefficients."""
        k = len(idx)
        pts = np.concatenate([self.P[idx].reshape(-1), np.frombuffer(self.C, dtype=np.uint8),
                              self.U[idx].reshape(-1), np.frombuffer(self.W, dtype=np.uint8)]).tobytes()
        sc = deltas + encode_fr(1)
        out = backend.msm_batch_encoded(pts, sc + sc, k + 1, 2)
        self.C, self.W = out[:64], out[64:128]
        for i, d in zip(idx.tolist(), dvals):
            cur = int.from_bytes(self.f[i].tobytes(), "little")
            self.f[i] = np.frombuffer(((cur + d) % R).to_bytes(32, "little"), dtype=np.uint8)
            self.y = (self.y + d * self.l[i]) % R
        self.evals = backend.ntt_encoded(self.f.tobytes(), self.n, False)
        return self.C + self.W + self.y.to_bytes(32, "little")

    def apply_verified(self, backend: HipBackend, idx: np.ndarray, deltas: bytes, dvals: List[int]) -> bytes:
        """apply, then the reference's per-update verification (dynaverse.rs:
        447-451) as one pairing-check call for this one proof."""
        from .pairing import verify_update_proofs
        proof = self.apply(backend, idx, deltas, dvals)
        if not verify_update_proofs(backend, [proof], [self.divisor])[0]:
            raise VerificationError(f"session {self.spec.sid}: update proof failed its pairing check")
        return proof


# ---- seeded request streams ------------------------------------------------------

SIZE_MIX = ((256, 0.4), (1024, 0.3), (4096, 0.2), (65536, 0.1))


def session_pool(per_size: int = 0, seed: int = 6, counts: Optional[Sequence[int]] = None) -> List[SessionSpec]:
    """Seeded sessions over the size mix: `per_size` of every size, or
    `counts[i]` of size SIZE_MIX[i]."""
    specs, sid = [], 0
    for j, (n, _) in enumerate(SIZE_MIX):
        for _ in range(counts[j] if counts is not None else per_size):
            specs.append(SessionSpec(sid, n, seed * 1_000_003 + sid))
            sid += 1
    return specs


def request_stream(specs: Sequence[SessionSpec], count: int, qps: float, seed: int = 6, k_max: int = 64):
    """Poisson arrivals at `qps`: (arrival offset s, sid, indices, deltas)."""
    rng = ChaCha20Stream(seed * 7919 + 1)
    by_size: Dict[int, List[SessionSpec]] = collections.defaultdict(list)
    for s in specs:
        by_size[s.n].append(s)
    cum, acc = [], 0.0
    for n, p in SIZE_MIX:
        acc += p
        cum.append((acc, n))
    t = 0.0
    out = []
    for _ in range(count):
        u = rng.below(1 << 30) / float(1 << 30)
        t += -np.log(1.0 - u) / qps if qps > 0 else 0.0
        v = rng.below(1 << 30) / float(1 << 30)
        n = next(nn for c, nn in cum if v < c or c == cum[-1][0])
        spec = by_size[n][rng.below(len(by_size[n]))]
        k = 1 + rng.below(k_max)
        idx = [rng.below(n) for _ in range(k)]
        ds = rng.uniform_fr(k)
        out.append((t, spec.sid, idx, ds))
    return out


def percentile(xs: Sequence[float], p: float) -> float:
    if not xs:
        return float("nan")
    s = sorted(xs)
    k = min(len(s) - 1, max(0, int(round(p / 100.0 * (len(s) - 1)))))
    return s[k]


def verify_served(backend: HipBackend, store: SessionStore, requests: Sequence[Request],
                  chunk: int = 1 << 16) -> Tuple[List[bool], float]:
    """Pairing verification of served requests' proofs on `backend` (the
    caller's owner, with the server closed): e(C' - [y'] g1, g2) == e(W', D)
    with each request's session divisor.  Returns (per-request results,
    wall ms)."""
    from .pairing import verify_update_proofs
    t0 = time.perf_counter()
    ok: List[bool] = []
    for i in range(0, len(requests), chunk):
        part = requests[i:i + chunk]
        ok.extend(verify_update_proofs(backend, [r.future.result() for r in part],
                                       [store.sessions[r.sid].divisor for r in part]))
    return ok, (time.perf_counter() - t0) * 1e3
