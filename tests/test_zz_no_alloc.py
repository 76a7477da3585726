"""Allocation-free batched MSM (dvgpu_ext_msm_batch_finalize_noalloc) against
the allocating entry point it is a variant of (dvgpu_ext_msm_batch_finalize).

Two calls in a row on one owner, the second smaller than the first (the
scratch array then holds the larger call's leftovers), for G1 and G2, with
an empty item and an indexed (sparse) item; every result must be byte-equal
to the allocating entry point's and to the oracle's closed form.  A batch
larger than the scratch array fails with -1 and leaves the owner usable.
"""
import numpy as np
import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.batching import MsmItem, msm_batch_finalize, msm_batch_finalize_noalloc
from dvgpu.codec import FR_MODULUS
from dvgpu.rng import ChaCha20Stream
from dvgpu.workload import tau_powers
from oracle import bn254 as o

pytestmark = pytest.mark.gpu

LOG_N = 16
N = 1 << LOG_N
R = FR_MODULUS


def _ints(arr):
    return [int.from_bytes(arr[i].tobytes(), "little") for i in range(arr.shape[0])]


@pytest.fixture(scope="module")
def env():
    b = HipBackend()
    tau = ChaCha20Stream(7101).nonzero_fr()
    pw = tau_powers(tau, N)
    d_pw = b.to_device(pw.reshape(-1))
    srs = b.alloc(64 * N)
    b.fixed_base_device(d_pw.ptr, N, srs.ptr)
    srs2 = b.alloc(128 * 4096)
    b.fixed_base_device(d_pw.ptr, 4096, srs2.ptr, g2=True)
    d_pw.free()
    sc = ChaCha20Stream(7102).uniform_fr_array(N)
    d_sc = b.to_device(sc.reshape(-1))
    idx = ChaCha20Stream(7103)
    index = np.array([idx.below(N) for _ in range(3000)], dtype=np.uint32)
    d_idx = b.to_device(index.view(np.uint8))
    yield b, _ints(pw), _ints(sc), index, srs, srs2, d_sc, d_idx
    b.close()


def _exponent(pw, sc, it, index=None):
    pts = (index[it.point_base:it.point_base + it.n] if index is not None
           else range(it.point_base, it.point_base + it.n))
    return sum(sc[it.scalar_base + j] * pw[int(p)] for j, p in enumerate(pts)) % R


def _compare(b, points, scalars, items, g2, index=None):
    width = 128 if g2 else 64
    out_a = b.alloc(width * len(items))
    out_b = b.alloc(width * len(items))
    want = msm_batch_finalize(b, points, scalars, items, g2=g2, index=index, results=out_a)
    got = msm_batch_finalize_noalloc(b, points, scalars, items, out_b, g2=g2, index=index)
    assert got == want
    assert out_b.download().tobytes() == out_a.download().tobytes()
    out_a.free()
    out_b.free()
    return got


def test_g1_two_calls_second_smaller(env):
    b, pw, sc, _, srs, _, d_sc, _ = env
    first = [MsmItem(N, 0, 0), MsmItem(0, 0, 0), MsmItem(4096, 1000, 20000), MsmItem(1000, 7, 7, window_bits=9)]
    second = [MsmItem(2048, 300, 5), MsmItem(777, 0, 60000)]
    for call, items in (("first", first), ("second, smaller", second)):
        print(f"case: no-alloc G1 {call} call, sizes {[it.n for it in items]}", flush=True)
        got = _compare(b, srs, d_sc, items, False)
        for it, r in zip(items, got):
            e = _exponent(pw, sc, it)
            assert r == o.g1_to_bytes(o.g1_mul(e) if e else None), it


def test_g1_indexed_points(env):
    b, pw, sc, index, srs, _, d_sc, d_idx = env
    items = [MsmItem(3000, 0, 100), MsmItem(500, 2000, 9)]
    print("case: no-alloc G1 indexed", flush=True)
    got = _compare(b, srs, d_sc, items, False, index=d_idx)
    for it, r in zip(items, got):
        assert r == o.g1_to_bytes(o.g1_mul(_exponent(pw, sc, it, index))), it


def test_g2_two_calls_second_smaller(env):
    b, pw, sc, _, _, srs2, d_sc, _ = env
    for call, items in (("first", [MsmItem(4096, 0, 0), MsmItem(1024, 512, 3000)]),
                        ("second, smaller", [MsmItem(300, 11, 17)])):
        print(f"case: no-alloc G2 {call} call", flush=True)
        got = _compare(b, srs2, d_sc, items, True)
        for it, r in zip(items, got):
            assert r == o.g2_to_bytes(o.g2_mul(_exponent(pw, sc, it))), it


def test_oversize_batch_rejected_owner_usable(env):
    b, pw, sc, _, srs, _, d_sc, _ = env
    out = b.alloc(64 * 8)
    print("case: no-alloc oversize batch", flush=True)
    with pytest.raises(BackendError) as e:
        msm_batch_finalize_noalloc(b, srs, d_sc, [MsmItem(N, 0, 0)] * 8, out)
    assert e.value.status == -1 and "scratch" in str(e.value)
    with pytest.raises(BackendError):                    # host-side range check
        msm_batch_finalize_noalloc(b, srs, d_sc, [MsmItem(N + 1, 0, 0)], out)
    it = MsmItem(5000, 123, 456)
    got = msm_batch_finalize_noalloc(b, srs, d_sc, [it], out)
    assert got == [o.g1_to_bytes(o.g1_mul(_exponent(pw, sc, it)))]
    out.free()
