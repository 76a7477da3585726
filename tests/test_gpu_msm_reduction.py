"""The single-MSM bucket reductions against the running-sum reduction and the
closed form.  Single MSMs (dvgpu_ext_msm_device, dvgpu_bn254_msm) sum each
window's buckets by bit planes of the bucket weights (small bucket sets, c <=
13) or by chunk running sums and bit planes of the chunk index (c >= 14).  The
batched paths keep the running sums, combined on the host
(dvgpu_ext_msm_batch_device) or on the device (dvgpu_ext_msm_batch_finalize).
All of them must equal sum_i s_i [tau^i] G = [f(tau)] G, computed by the
oracle on the host (Horner over Fr, one scalar multiplication), on both
sides of the mode switch, for uniform, sparse and zero scalars."""
import numpy as np
import pytest

from dvgpu import HipBackend
from dvgpu.batching import MsmItem, msm_batch_device, msm_batch_finalize
from dvgpu.codec import FR_MODULUS
from dvgpu.rng import ChaCha20Stream
from dvgpu.workload import poly_eval, srs_scalars
from oracle import bn254 as o

pytestmark = pytest.mark.gpu

N = 1 << 16      # SRS points (G1)
N2 = 1024        # G2 points


@pytest.fixture(scope="module")
def env():
    b = HipBackend()
    tau, pw = srs_scalars(4101, N)
    d_pw = b.to_device(pw.reshape(-1))
    g1 = b.alloc(64 * N)
    b.fixed_base_device(d_pw.ptr, N, g1.ptr)
    g2 = b.alloc(128 * N2)
    b.fixed_base_device(d_pw.ptr, N2, g2.ptr, g2=True)
    d_pw.free()
    b.powers = pw
    yield b, tau, g1, g2
    b.close()


def scalars(kind, n, seed):
    if kind == "uniform":
        return ChaCha20Stream(seed).uniform_fr_array(n)
    out = np.zeros((n, 32), dtype=np.uint8)
    if kind == "sparse":
        # a few values: most buckets empty, a few very full; -1 puts the top
        # digit of every window at its extreme
        pick = ChaCha20Stream(seed).uniform_fr_array(n)[:, 0] % 6
        vals = [0, 1, 2, 3, 1 << 40, FR_MODULUS - 1]
        for i in range(n):
            out[i] = np.frombuffer(vals[pick[i]].to_bytes(32, "little"), dtype=np.uint8)
    return out


def expect(b, sc, tau, g2):
    """[f(tau)] G from the oracle alone (no device computation)."""
    f = poly_eval(sc, tau)
    return o.g2_to_bytes(o.g2_mul(f)) if g2 else o.g1_to_bytes(o.g1_mul(f))


@pytest.mark.parametrize("kind", ["uniform", "sparse", "zero"])
@pytest.mark.parametrize("c", [4, 9, 13, 14, 16])
def test_g1_reductions_agree(env, c, kind):
    b, tau, g1, _ = env
    n = 1 << 13
    sc = scalars(kind, n, 11 + c)
    d_sc = b.to_device(sc.reshape(-1))
    single = b.msm_device(g1.ptr, d_sc.ptr, n, window_bits=c)
    batch = msm_batch_device(b, g1, d_sc, [MsmItem(n, 0, 0, c)])[0]
    fin = msm_batch_finalize(b, g1, d_sc, [MsmItem(n, 0, 0, c)])[0]
    want = expect(b, sc, tau, False)
    assert single == batch == fin == want, (c, kind)
    if kind == "zero":
        assert single == bytes(64)
    d_sc.free()


def test_g1_default_windows_at_2p16(env):
    """The 2^16 default (c = 13, bucket-weight planes) and a forced c = 15
    (chunk planes) over the whole SRS, through the plain ABI as well."""
    b, tau, g1, _ = env
    sc = scalars("uniform", N, 77)
    d_sc = b.to_device(sc.reshape(-1))
    want = expect(b, sc, tau, False)
    assert b.msm_device(g1.ptr, d_sc.ptr, N) == want
    assert b.msm_device(g1.ptr, d_sc.ptr, N, window_bits=15) == want
    pts = b"".join(b.fixed_base_encoded(b.powers[i:i + 4096].tobytes(), 4096, False) for i in range(0, N, 4096))
    assert b.msm_encoded(pts, sc.tobytes(), N) == want      # dvgpu_bn254_msm
    d_sc.free()


@pytest.mark.parametrize("c", [6, 13, 14])
def test_g2_reductions_agree(env, c):
    b, tau, _, g2 = env
    sc = scalars("uniform", N2, 300 + c)
    d_sc = b.to_device(sc.reshape(-1))
    single = b.msm_device(g2.ptr, d_sc.ptr, N2, g2=True, window_bits=c)
    batch = msm_batch_device(b, g2, d_sc, [MsmItem(N2, 0, 0, c)], g2=True)[0]
    assert single == batch == expect(b, sc, tau, True), c
    d_sc.free()
