"""The single-call drop-in workload at its full size against the oracle.

A BN254 G1 MSM of 2^20 points over the seeded SRS P_i = [tau^i] g1 with
uniform Fr scalars s_i (the benchmark's own seeds) must equal the oracle's
closed form [f(tau)] g1, f(X) = sum s_i X^i: Horner over Fr on the host, then
one oracle scalar multiplication.  Nothing on the expected side runs on the
device, so a wrong SRS (fixed-base kernel) fails as well as a wrong MSM.

Covered through the device entry point (dvgpu_ext_msm_device, the timed path)
and the plain ABI (dvgpu_bn254_msm, host canonical points and scalars), at
2^20 and at the 2^16 prefix (the configuration's second size), with the
default window and one forced other width.  Reference: the n-sweep of
crates/gpu/tests/icicle_backend.rs:64-116, which compares against arkworks.
"""
import numpy as np
import pytest

from dvgpu import HipBackend
from dvgpu.rng import ChaCha20Stream
from dvgpu.workload import poly_eval, srs_scalars
from oracle import bn254 as o

pytestmark = pytest.mark.gpu

LOG_N = 20
N = 1 << LOG_N


@pytest.fixture(scope="module")
def env():
    b = HipBackend()
    tau, powers = srs_scalars(2, N)
    scal = ChaCha20Stream(3).uniform_fr_array(N)
    d_pow = b.to_device(powers.reshape(-1))
    d_srs = b.alloc(64 * N)
    b.fixed_base_device(d_pow.ptr, N, d_srs.ptr)
    d_pow.free()
    d_sc = b.to_device(scal.reshape(-1))
    yield b, tau, powers, scal, d_srs, d_sc
    b.close()


def oracle_commitment(scal, tau):
    """[sum_i s_i tau^i] g1 in canonical affine bytes, all on the host."""
    return o.g1_to_bytes(o.g1_mul(poly_eval(scal, tau)))


def canonical_srs(b, powers, n):
    """Canonical affine SRS points for the plain ABI, through the setup ABI."""
    return b"".join(b.fixed_base_encoded(powers[i:i + 4096].tobytes(), min(4096, n - i), False)
                    for i in range(0, n, 4096))


def test_srs_points_match_oracle(env):
    """Sampled SRS points equal [tau^i] g1 from the oracle (the closed form's premise)."""
    b, tau, powers, _, _, _ = env
    for i in (0, 1, 2, 4095, 4096, 65535, 65536, N - 1):
        got = b.fixed_base_encoded(powers[i].tobytes(), 1, False)
        assert got == o.g1_to_bytes(o.g1_mul(pow(tau, i, o.R))), i


@pytest.mark.parametrize("window", [0, 12, 13, 14, 16])
def test_msm_2p20_device_entry(env, window):
    """The default (c = 15, chunk-plane reduction), narrower windows (bucket-
    or chunk-plane reductions, more windows) and c = 16 (2^15 buckets: the
    global-memory digit path)."""
    b, tau, _, scal, d_srs, d_sc = env
    print(f"2^20 G1 MSM, device entry, window {window or 'default'}", flush=True)
    assert b.msm_device(d_srs.ptr, d_sc.ptr, N, window_bits=window) == oracle_commitment(scal, tau)


def test_msm_2p20_plain_abi(env):
    b, tau, powers, scal, _, _ = env
    print("2^20 G1 MSM, plain ABI (host buffers)", flush=True)
    pts = canonical_srs(b, powers, N)
    assert b.msm_encoded(pts, scal.tobytes(), N) == oracle_commitment(scal, tau)


def test_msm_2p16_both_entries(env):
    b, tau, powers, scal, d_srs, d_sc = env
    n = 1 << 16
    want = oracle_commitment(scal[:n], tau)
    assert b.msm_device(d_srs.ptr, d_sc.ptr, n) == want
    assert b.msm_encoded(canonical_srs(b, powers, n), scal[:n].tobytes(), n) == want


def test_msm_2p20_linearity(env):
    """Size-independent check at 2^20: MSM(s) + MSM(t) == MSM(s + t) (group
    addition in the oracle), with t a second uniform vector."""
    b, _, _, scal, d_srs, d_sc = env
    t = ChaCha20Stream(33).uniform_fr_array(N)
    st = np.frombuffer(b"".join(((int.from_bytes(x.tobytes(), "little") + int.from_bytes(y.tobytes(), "little"))
                                 % o.R).to_bytes(32, "little") for x, y in zip(scal, t)),
                       dtype=np.uint8).reshape(N, 32)
    d_t = b.to_device(t.reshape(-1))
    d_st = b.to_device(st.reshape(-1))
    a = o.g1_from_bytes(b.msm_device(d_srs.ptr, d_sc.ptr, N))
    c = o.g1_from_bytes(b.msm_device(d_srs.ptr, d_t.ptr, N))
    s = o.g1_from_bytes(b.msm_device(d_srs.ptr, d_st.ptr, N))
    assert o.add(a, c) == s
    d_t.free()
    d_st.free()


@pytest.mark.parametrize("log_n", [17, 18, 19])
def test_msm_prefix_sizes(env, log_n):
    """2^17, 2^18 and 2^19 (their default window widths) over the SRS prefix,
    against the oracle's closed form."""
    b, tau, _, scal, d_srs, d_sc = env
    n = 1 << log_n
    print(f"G1 MSM 2^{log_n}, device entry", flush=True)
    assert b.msm_device(d_srs.ptr, d_sc.ptr, n) == oracle_commitment(scal[:n], tau)


@pytest.mark.parametrize("window", [0, 14])
def test_msm_2p20_one_scalar_everywhere(env, window):
    """Every scalar equal: in every window all 2^20 entries fall into one
    bucket, so one placement group holds them all and every staging block
    sends its whole chunk (2^15 points at c = 15, 2^14 at c = 14) to that
    group over several LDS rounds, and the group's placement runs 256 rounds.
    Expected: [s * sum_i tau^i] g1 by the geometric sum, on the host."""
    b, tau, _, _, d_srs, _ = env
    s = ChaCha20Stream(9).nonzero_fr()
    d = b.to_device(np.frombuffer(s.to_bytes(32, "little") * N, dtype=np.uint8))
    geo = (pow(tau, N, o.R) - 1) * pow(tau - 1, o.R - 2, o.R) % o.R
    try:
        assert b.msm_device(d_srs.ptr, d.ptr, N, window_bits=window) == o.g1_to_bytes(o.g1_mul(s * geo % o.R))
    finally:
        d.free()
