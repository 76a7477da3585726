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
from dvgpu.batching import DeviceBatch, MsmItem, NttRows, msm_batch_device, set_window_table
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


def test_window_table_changes_plan_not_results(backend, wl):
    table = [0] * 25
    for lg in range(10, 19):
        table[lg] = max(4, lg - 6)
    set_window_table(backend, table)
    try:
        print("case: window table override", flush=True)
        assert msm_batch_device(backend, wl.srs, wl.d_sc, wl.items[:16]) == wl.exp1[:16]
    finally:
        set_window_table(backend, [0] * 25)
    with pytest.raises(BackendError):
        set_window_table(backend, [3] * 25)


def test_invalid_batches_leave_owner_healthy(backend, wl):
    with pytest.raises(BackendError):
        msm_batch_device(backend, wl.srs, wl.d_sc, [MsmItem(16, 0, 0, window_bits=3)])
    with pytest.raises(BackendError):
        DeviceBatch(backend, ntt=[NttRows(wl.in12, wl.out12, 3000, 1)])
    assert msm_batch_device(backend, wl.srs, wl.d_sc, [MsmItem(0, 0, 0)]) == [bytes(64)]
    assert msm_batch_device(backend, wl.srs, wl.d_sc, wl.items[:2]) == wl.exp1[:2]
