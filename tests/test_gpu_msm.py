"""G1 / G2 MSM on the MI355X through the C ABI, bit-exact against the oracle.

Cases follow the reference's own tests: crates/gpu/tests/icicle_backend.rs
(sizes 0..8192, identity points every 7th, zero scalars every 11th,
Sigma i^2 G, 3n G, 5G, G - G = O, all-zero scalars), sparse_msm_backend.rs and
batch offsets.  Large sizes are checked through the discrete-log closed form.
"""
import pytest

from dvgpu import BackendError, BatchMsmRequest, HipBackend, SparseMsmRequest
from dvgpu.rng import ChaCha20Stream
from oracle import bn254 as o
from tests.helpers import closed_form_msm, progression_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


def test_identity_and_small_known_answers(backend):
    G = o.G1_GEN
    print("case: known answers", flush=True)
    assert backend.msm([], []) is None
    assert backend.msm([G], [3]) == o.g1_mul(3)
    assert backend.msm([G, G], [1, o.R - 1]) is None                    # icicle_backend.rs:89-91
    assert backend.msm([G, G, None], [2, 3, 99]) == o.g1_mul(5)         # icicle_backend.rs:162-166
    assert backend.msm([G] * 5, [0] * 5) is None                       # all-zero scalars
    n = 64
    pts = [o.g1_mul(i) for i in range(1, n + 1)]
    assert backend.msm(pts, list(range(1, n + 1))) == o.g1_mul(sum(i * i for i in range(1, n + 1)))


@pytest.mark.parametrize("n", [1024, 8192])
def test_repeated_point_hot_bucket(backend, n):
    # icicle_backend.rs:124-128: n copies of G with scalar 3 -> [3n] G.  Every
    # point lands in one bucket, which exercises the bucket-splitting tasks.
    print(f"case: hot bucket n={n}", flush=True)
    G = o.G1_GEN
    assert backend.msm([G] * n, [3] * n) == o.g1_mul(3 * n)


@pytest.mark.parametrize("seed", [7, 17, 2026])
def test_sizes_with_identity_and_zero_patterns(backend, seed):
    rng = ChaCha20Stream(seed)
    a, b = rng.nonzero_fr(), rng.nonzero_fr()
    base_pts, logs = progression_points(8192, a, b)
    for n in [1, 2, 17, 511, 512, 1022, 1023, 1024, 2048, 6144, 8192]:
        print(f"case: seed={seed} n={n}", flush=True)
        pts = [None if i % 7 == 0 else base_pts[i] for i in range(n)]
        scalars = rng.uniform_fr(n)
        scalars = [0 if i % 11 == 0 else s for i, s in enumerate(scalars)]
        exp = closed_form_msm(logs[:n], scalars, mask_identity=range(0, n, 7))
        assert backend.msm(pts, scalars) == exp, (seed, n)
        assert backend.msm(pts, [0] * n) is None


def test_straight_oracle_small_random(backend):
    rng = ChaCha20Stream(99)
    for n in (3, 40):
        pts = [o.g1_mul(k) for k in rng.uniform_fr(n)]
        sc = rng.uniform_fr(n)
        assert backend.msm(pts, sc) == o.straight_msm(pts, sc)


def test_window_sizes_agree(backend):
    # Every window width gives the same group element (c = 4 .. 16).
    rng = ChaCha20Stream(5)
    pts, logs = progression_points(3000, rng.nonzero_fr(), rng.nonzero_fr())
    sc = rng.uniform_fr(3000)
    enc_p = b"".join(o.g1_to_bytes(p) for p in pts)
    enc_s = b"".join(o.fr_to_bytes(s) for s in sc)
    exp = o.g1_to_bytes(closed_form_msm(logs, sc))
    dev_p = backend.to_device(enc_p)
    dev_s = backend.to_device(enc_s)
    mont = backend.alloc(len(enc_p))
    backend.points_to_mont_device(dev_p.ptr, mont.ptr, 3000)
    for c in (4, 5, 7, 8, 10, 11, 12, 13, 14, 15, 16):
        print(f"case: window c={c}", flush=True)
        assert backend.msm_device(mont.ptr, dev_s.ptr, 3000, window_bits=c) == exp, c


def test_batch_offsets_and_rejections(backend):
    # icicle_backend.rs:203-225
    G = o.G1_GEN
    req = BatchMsmRequest(point_sets=[[G, G], [G]], shared_scalars=[9, 2, 3], offsets=[1, 0], lengths=[2, 1])
    assert backend.msm_batch(req) == [o.g1_mul(5), o.g1_mul(9)]
    for offsets, lengths in (([2, 0], [2, 1]), ([0], [2, 1]), ([0, 0], [3, 1])):
        with pytest.raises(ValueError):
            backend.msm_batch(BatchMsmRequest([[G, G], [G]], [9, 2, 3], offsets, lengths))
    with pytest.raises(ValueError):
        backend.msm([(1, 1)], [1])          # off-curve point rejected before native
    assert backend.msm([G], [1]) == G       # owner still healthy


def test_sparse_batch_repeats_and_bins(backend):
    # sparse_msm_backend.rs:71-76: repeated index 70 x 17 x 5 -> [85] G style
    G = o.G1_GEN
    bases, logs = progression_points(300, 1, 1)   # bases[i] = [i+1] G
    reqs = [
        SparseMsmRequest(bases, [(0, 17)] * 5),                           # [85] G
        SparseMsmRequest(bases, [(i, i + 3) for i in range(0, 300, 7)]),
        SparseMsmRequest(bases, []),
        SparseMsmRequest(bases, [(299, o.R - 1), (0, 5)]),
    ]
    got = backend.msm_sparse_batch(reqs)
    assert got[0] == o.g1_mul(85)
    assert got[1] == o.g1_mul(sum((i + 3) * (i + 1) for i in range(0, 300, 7)))
    assert got[2] is None
    assert got[3] == o.g1_mul((o.R - 1) * 300 + 5)
    with pytest.raises(ValueError):
        backend.msm_sparse_batch([SparseMsmRequest(bases, [(300, 1)])])


def test_batch_many_sets_uneven(backend):
    rng = ChaCha20Stream(44)
    pts, logs = progression_points(700, rng.nonzero_fr(), rng.nonzero_fr())
    lengths = [1, 5, 64, 65, 300, 700, 3, 128]
    sc = rng.uniform_fr(sum(lengths))
    offsets, o_ = [], 0
    for ln in lengths:
        offsets.append(o_)
        o_ += ln
    req = BatchMsmRequest([pts] * len(lengths), sc, offsets, lengths)
    got = backend.msm_batch(req)
    for i, ln in enumerate(lengths):
        assert got[i] == closed_form_msm(logs[:ln], sc[offsets[i]:offsets[i] + ln]), i
    st = backend.batch_stats()
    assert st["icicle_calls"] >= 1 and st["max_batch_size"] >= 1


def test_g2_msm_matches_oracle(backend):
    g2 = o.G2_GEN
    rng = ChaCha20Stream(8)
    ks = rng.uniform_fr(6)
    pts = [o.g2_mul(k) for k in ks] + [None]
    sc = rng.uniform_fr(7)
    exp = o.mul(o.G2_GEN, sum(s * k for s, k in zip(sc, ks)) % o.R, o.FQ2)
    assert backend.msm_g2_host(pts, sc) == exp
    assert backend.msm_g2_host([g2, g2], [1, o.R - 1]) is None
    assert backend.msm_g2_host([g2] * 300, [3] * 300) == o.g2_mul(900)


def test_invalid_arguments_do_not_retire(backend):
    with pytest.raises(BackendError) as e:
        backend.msm_batch_encoded(b"\0" * 64, b"\0" * 32, 1, 65)
    assert e.value.status == -1
    assert backend.msm([o.G1_GEN], [2]) == o.g1_mul(2)


def test_large_factor_owner_and_identity():
    with pytest.raises(ValueError):
        HipBackend(101)
    b = HipBackend(10)
    assert "gfx950" in b.device_identity()
    free, total = b.memory_info()
    assert 0 < free <= total
    assert b.msm([o.G1_GEN] * 4, [1, 2, 3, 4]) == o.g1_mul(10)
    b.close()
