"""Batching form on the MI355X: many uneven MSMs (G1 and G2), NTT row batches
and G2 subgroup validation in one launch sequence, eager and as a hipGraph.

Workload of the batched configuration: 64 G1 MSMs with sizes log-uniform in
[2^10, 2^18] over a device SRS [tau^i] g1, 8 G2 MSMs of 2^10 over
[tau^i] g2, 16 NTT rows of 2^12 and 16 of 2^16, 4096 G2 points with every odd
one pushed into a non-subgroup coset.  Every MSM is checked in closed form
([sum s_j tau^(p+j)] G, evaluated by the oracle), every NTT row against the C
oracle, masks against the corpus construction.  A graph is replayed several
times with eager work between the replays and every replay is checked.
"""
import numpy as np
import pytest

from dvgpu import BackendError, HipBackend
from dvgpu import autotune
from dvgpu.batching import DeviceBatch, MsmItem, NttRows, msm_batch_device, msm_batch_finalize, ntt_batch_device
from dvgpu.codec import FR_MODULUS, encode_fr, root_of_unity_2_20
from dvgpu.rng import ChaCha20Stream
from dvgpu.workload import tau_powers
from oracle import bn254 as o
from oracle import cref
from tests.g2corpus import twist_cofactor_point

pytestmark = pytest.mark.gpu

LOG_SRS = 18
R = FR_MODULUS


def _ints(arr):
    return [int.from_bytes(arr[i].tobytes(), "little") for i in range(arr.shape[0])]


class Workload:
    def __init__(self, backend):
        b = self.b = backend
        n = 1 << LOG_SRS
        rng = ChaCha20Stream(3003)
        self.tau = rng.nonzero_fr()
        pw = tau_powers(self.tau, n)
        self.pw = _ints(pw)
        d_pw = b.to_device(pw.reshape(-1))
        self.srs = b.alloc(64 * n)
        b.fixed_base_device(d_pw.ptr, n, self.srs.ptr)
        self.srs2 = b.alloc(128 * 1024)
        b.fixed_base_device(d_pw.ptr, 1024, self.srs2.ptr, g2=True)
        d_pw.free()
        sc = ChaCha20Stream(3004).uniform_fr_array(n)
        self.sc = _ints(sc)
        self.d_sc = b.to_device(sc.reshape(-1))
        # prefix[k] = sum_{i<k} s_i tau^i: items with scalar_base == point_base
        # have exponent prefix[p+n] - prefix[p].
        self.prefix = [0] * (n + 1)
        acc = 0
        for i in range(n):
            acc = (acc + self.sc[i] * self.pw[i]) % R
            self.prefix[i + 1] = acc
        # 64 G1 items, log-uniform sizes; every 4th has its own scalar offset.
        u = ChaCha20Stream(3005)
        self.items = []
        for i in range(64):
            ln = 10 + 8 * (u.below(1 << 20) / (1 << 20))
            m = max(1 << 10, min(n, int(round(2 ** ln))))
            pb = u.below(n - m + 1)
            sb = pb if i % 4 else u.below(n - m + 1)
            if i % 4 == 3:
                m = min(m, 1 << 13)
            self.items.append(MsmItem(m, pb, sb))
        self.items[5] = MsmItem(0, 0, 0)                 # empty set -> identity
        self.items[9] = MsmItem(1 << 18, 0, 0)            # the full SRS
        self.items[11].window_bits = 9                    # explicit width
        self.exp1 = [self.expect(it) for it in self.items]
        # 8 G2 MSMs of 2^10 at different scalar offsets.
        self.items2 = [MsmItem(1024, 0, 4096 * k) for k in range(8)]
        self.exp2 = [o.g2_to_bytes(o.g2_mul(self.exponent(it))) for it in self.items2]
        # NTT rows: 16 x 2^12 forward, 16 x 2^16 forward, then the 2^12 outputs
        # transformed back (a later entry reads an earlier entry's output).
        self.root = encode_fr(root_of_unity_2_20())
        v12 = ChaCha20Stream(3006).uniform_fr_array(16 << 12)
        v16 = ChaCha20Stream(3007).uniform_fr_array(16 << 16)
        self.v12, self.v16 = v12.reshape(-1), v16.reshape(-1)
        self.in12, self.in16 = b.to_device(self.v12), b.to_device(self.v16)
        self.out12, self.out16, self.back12 = b.alloc(32 << 16), b.alloc(32 << 20), b.alloc(32 << 16)
        self.ntt = [NttRows(self.in12, self.out12, 1 << 12, 16), NttRows(self.in16, self.out16, 1 << 16, 16),
                    NttRows(self.out12, self.back12, 1 << 12, 16, inverse=True)]
        self.exp12 = b"".join(cref.ntt(self.v12[(32 << 12) * r:(32 << 12) * (r + 1)].tobytes(), 1 << 12, False,
                                       self.root, 8) for r in range(16))
        self.exp16 = b"".join(cref.ntt(self.v16[(32 << 16) * r:(32 << 16) * (r + 1)].tobytes(), 1 << 16, False,
                                       self.root, 8) for r in range(16))
        # 4096 G2 points, odd ones in a non-subgroup coset.
        g2sc = ChaCha20Stream(3008).uniform_fr(4096)
        pts = b.fixed_base_g2(g2sc)
        T = twist_cofactor_point()
        wire = b"".join(o.g2_to_bytes(o.add(p, T, o.FQ2) if i % 2 else p) for i, p in enumerate(pts))
        self.wire = b.to_device(wire)
        self.masks = b.alloc(4 * 4096)
        self.exp_masks = np.array([4 if i % 2 else 0 for i in range(4096)], dtype=np.uint32)

    def exponent(self, it):
        if it.n == 0:
            return 0
        if it.point_base == it.scalar_base:
            return (self.prefix[it.point_base + it.n] - self.prefix[it.point_base]) % R
        return sum(self.sc[it.scalar_base + j] * self.pw[it.point_base + j] for j in range(it.n)) % R

    def expect(self, it):
        e = self.exponent(it)
        return o.g1_to_bytes(o.g1_mul(e) if e else None)


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


@pytest.fixture(scope="module")
def wl(backend):
    print("case: building batched workload", flush=True)
    return Workload(backend)


def _check_batch_outputs(wl, r1, r2, rep):
    assert r1 == wl.exp1
    assert r2 == wl.exp2
    assert wl.out12.download().tobytes() == wl.exp12
    assert wl.out16.download().tobytes() == wl.exp16
    assert wl.back12.download().tobytes() == wl.v12.tobytes()
    masks = wl.masks.download().view(np.uint32)
    assert np.array_equal(masks, wl.exp_masks)
    assert rep["checked_points"] == 4096 and rep["invalid_points"] == 2048
    assert rep["first_invalid"] == 1 and rep["failure_mask"] == 4 and rep["algorithm"] == 3


def test_msm_batch_device_g1_uneven(backend, wl):
    print("case: msm_batch_device 64 uneven G1", flush=True)
    got, ph = msm_batch_device(backend, wl.srs, wl.d_sc, wl.items, phases=True)
    assert got == wl.exp1
    assert len(ph) == 7 and all(p >= 0 for p in ph)


def test_msm_batch_device_g2(backend, wl):
    print("case: msm_batch_device 8 G2", flush=True)
    assert msm_batch_device(backend, wl.srs2, wl.d_sc, wl.items2, g2=True) == wl.exp2


def test_each_result_same_alone_and_in_batch(backend, wl):
    # The batched result of an item equals the single-MSM entry point's.
    for k in (0, 9, 11, 17, 63):
        it = wl.items[k]
        if it.n == 0:
            continue
        print(f"case: alone item {k} n={it.n}", flush=True)
        alone = backend.msm_device(wl.srs.ptr + 64 * it.point_base, wl.d_sc.ptr + 32 * it.scalar_base, it.n,
                                   window_bits=it.window_bits)
        assert alone == wl.exp1[k]


def test_sparse_index_batch(backend, wl):
    # Sparse update sets: items address the SRS through an index array, with
    # repeated indices, k in {1, 16, 256, 4096}.
    u = ChaCha20Stream(3010)
    n = 1 << LOG_SRS
    ks = [1, 16, 256, 4096, 4096, 256, 16, 1]
    hot = [u.below(n) for _ in range(n // 100)]
    idx, items, exps, pos = [], [], [], 0
    for j, k in enumerate(ks):
        sel = [hot[u.below(len(hot))] if j % 2 else u.below(n) for _ in range(k)]
        sb = u.below(n - k + 1)
        items.append(MsmItem(k, pos, sb))
        e = sum(wl.sc[sb + t] * wl.pw[p] for t, p in enumerate(sel)) % R
        exps.append(o.g1_to_bytes(o.g1_mul(e) if e else None))
        idx.extend(sel)
        pos += k
    d_idx = backend.to_device(np.array(idx, dtype=np.uint32).view(np.uint8))
    print("case: sparse index batch", flush=True)
    assert msm_batch_device(backend, wl.srs, wl.d_sc, items, index=d_idx) == exps


@pytest.mark.parametrize("use_graph", [False, True])
def test_full_batch_eager_and_graph_replays(backend, wl, use_graph):
    print(f"case: full batch use_graph={use_graph}", flush=True)
    batch = DeviceBatch(backend, g1=(wl.srs, wl.d_sc, wl.items), g2=(wl.srs2, wl.d_sc, wl.items2), ntt=wl.ntt,
                        g2_wire=wl.wire, g2_wire_n=4096, g2_algorithm=3, g2_masks=wl.masks, use_graph=use_graph)
    try:
        # The create call already ran the sequence once (warm-up).
        for rep_i in range(3):
            print(f"case: replay {rep_i}", flush=True)
            # Eager work between replays, touching the same output buffers.
            wl.out12.upload(np.zeros(32 << 16, dtype=np.uint8))
            wl.masks.upload(np.full(4 * 4096, 0xFF, dtype=np.uint8))
            other = backend.msm_device(wl.srs.ptr, wl.d_sc.ptr, 3000 + rep_i)
            assert other == wl.expect(MsmItem(3000 + rep_i, 0, 0))
            r1, r2, rep, ms = batch.run()
            assert ms > 0
            _check_batch_outputs(wl, r1, r2, rep)
    finally:
        batch.close()


def test_tuning_changes_plan_not_results(backend, wl):
    base = autotune.get(backend)
    table = {"msm_window_bits": {lg: max(4, lg - 6) for lg in range(10, 19)}, "msm_task_points": 7,
             "msm_chunk": 2, "ntt_pass_stages": {12: 4, 16: 6}}
    autotune.apply(backend, table)
    try:
        got = autotune.get(backend)
        assert got["msm_task_points"] == 7 and got["msm_chunk"] == 2 and got["ntt_pass_stages"] == {12: 4, 16: 6}
        print("case: tuning override msm", flush=True)
        assert msm_batch_device(backend, wl.srs, wl.d_sc, wl.items[:16]) == wl.exp1[:16]
        print("case: tuning override ntt", flush=True)
        for length, d_in, d_out, exp in ((1 << 12, wl.in12, wl.out12, wl.exp12), (1 << 16, wl.in16, wl.out16, wl.exp16)):
            backend.ntt_device(d_in.ptr, d_out.ptr, length, 16, False)
            assert d_out.download().tobytes() == exp
    finally:
        autotune.apply(backend, base)
    for bad in ({"msm_window_bits": {12: 3}}, {"ntt_pass_stages": {12: 12}}, {"msm_chunk": 3},
                {"msm_task_points": 5000}):
        with pytest.raises(BackendError):
            autotune.apply(backend, bad)
    assert autotune.get(backend) == base


def test_invalid_batches_leave_owner_healthy(backend, wl):
    with pytest.raises(BackendError):
        msm_batch_device(backend, wl.srs, wl.d_sc, [MsmItem(16, 0, 0, window_bits=3)])
    with pytest.raises(BackendError):
        DeviceBatch(backend, ntt=[NttRows(wl.in12, wl.out12, 3000, 1)])
    assert msm_batch_device(backend, wl.srs, wl.d_sc, [MsmItem(0, 0, 0)]) == [bytes(64)]
    assert msm_batch_device(backend, wl.srs, wl.d_sc, wl.items[:2]) == wl.exp1[:2]


def test_fused_ntt_batch_mixed_lengths(backend):
    # Lengths 1 .. 2^16 with several row counts, forward and inverse, in one
    # fused launch sequence; each row against the C oracle.
    root = encode_fr(root_of_unity_2_20())
    shapes = [(1, 3, False), (2, 1, True), (256, 5, False), (1024, 2, True), (2048, 1, False), (4096, 3, False),
              (1 << 16, 2, True), (8, 7, True)]
    rng = ChaCha20Stream(3020)
    rows, host = [], []
    for length, r, inv in shapes:
        vals = rng.uniform_fr_array(length * r).reshape(-1)
        d_in = backend.to_device(vals)
        d_out = backend.alloc(vals.nbytes)
        rows.append(NttRows(d_in, d_out, length, r, inv))
        host.append((vals, length, r, inv))
    in_place = rng.uniform_fr_array(512 * 2).reshape(-1)
    d_ip = backend.to_device(in_place)
    rows.append(NttRows(d_ip, d_ip, 512, 2, False))
    host.append((in_place, 512, 2, False))
    print("case: fused ntt batch", flush=True)
    ms = ntt_batch_device(backend, rows, timed=True)
    assert ms > 0
    for (vals, length, r, inv), row in zip(host, rows):
        exp = b"".join(cref.ntt(vals[32 * length * k:32 * length * (k + 1)].tobytes(), length, inv, root, 8)
                       for k in range(r))
        assert row.out.download().tobytes() == exp, (length, r, inv)
    # Overlapping entries are rejected before any work.
    with pytest.raises(BackendError):
        ntt_batch_device(backend, [NttRows(rows[2].inp, rows[2].out, 256, 1), NttRows(rows[2].out, rows[3].out, 256, 1)])


def test_finalize_on_device_and_result_slots(backend, wl):
    # Device-side window sums equal the host epilogue; results stored into
    # slots replace inputs only after every MSM of the call has read them.
    items = wl.items[:24]
    print("case: finalize vs host epilogue", flush=True)
    assert msm_batch_finalize(backend, wl.srs, wl.d_sc, items) == wl.exp1[:24]
    # Commitment update in place: C' = C + sum d_j P_j, where C sits in a slot
    # array and enters the MSM through the index with scalar 1.
    u = ChaCha20Stream(3021)
    base = backend.alloc(64 * 8)
    # Points 0..3: P_0..P_3 (Montgomery, resident); 4..7: cached points, identity at first.
    base.upload(np.concatenate([wl.srs.download(64 * 4), np.zeros(64 * 4, dtype=np.uint8)]))
    ks = [3, 1, 5, 2]
    state = [0] * len(ks)                               # discrete log of each cached point
    for rnd in range(3):
        idx, sc, its, slots, exps = [], b"", [], [], []
        for j, k in enumerate(ks):
            sel = [u.below(4) for _ in range(k)]
            ds = [u.nonzero_fr() for _ in range(k)]
            its.append(MsmItem(k + 1, len(idx), len(sc) // 32))
            idx.extend(sel + [4 + j])
            sc += b"".join(encode_fr(d) for d in ds) + encode_fr(1)
            slots.append(4 + j)
            state[j] = (state[j] + sum(d * wl.pw[p] for d, p in zip(ds, sel))) % R
            exps.append(o.g1_to_bytes(o.g1_mul(state[j]) if state[j] else None))
        d_idx = backend.to_device(np.array(idx, dtype=np.uint32).view(np.uint8))
        d_sc = backend.to_device(sc)
        print(f"case: slot round {rnd}", flush=True)
        got = msm_batch_finalize(backend, base, d_sc, its, index=d_idx, slots=slots, slot_points=base)
        assert got == exps, rnd
    with pytest.raises(BackendError):                   # an empty MSM may not name a slot
        msm_batch_finalize(backend, base, d_sc, [MsmItem(0, 0, 0)], index=d_idx, slots=[4], slot_points=base)


def test_finalize_rejects_out_of_range_points(backend, wl):
    base = backend.alloc(64 * 8)
    base.upload(wl.srs.download(64 * 8))
    before = base.download().tobytes()
    sc = backend.to_device(encode_fr(5) * 4)
    bad = backend.to_device(np.array([1, 2, 8, 3], dtype=np.uint32).view(np.uint8))   # 8 is out of range
    print("case: finalize index out of range", flush=True)
    with pytest.raises(BackendError):
        msm_batch_finalize(backend, base, sc, [MsmItem(4, 0, 0)], index=bad, slots=[7], slot_points=base)
    assert base.download().tobytes() == before          # no slot written
    with pytest.raises(BackendError):                   # host-side range checks
        msm_batch_finalize(backend, base, sc, [MsmItem(5, 0, 0)])
    with pytest.raises(BackendError):
        msm_batch_finalize(backend, base, sc, [MsmItem(2, 0, 0)], slots=[8], slot_points=base)
    good = backend.to_device(np.array([1, 2, 7, 3], dtype=np.uint32).view(np.uint8))
    e = 5 * (wl.pw[1] + wl.pw[2] + wl.pw[7] + wl.pw[3]) % R
    assert msm_batch_finalize(backend, base, sc, [MsmItem(4, 0, 0)], index=good) == [o.g1_to_bytes(o.g1_mul(e))]
