"""Fr NTT/INTT on the MI355X through the C ABI, bit-exact against the oracle.

Cases follow crates/gpu/tests/ntt_backend.rs: sizes 1..16384 forward and
inverse against the CPU transform, round trips, impulse -> ones, constant c ->
(n c, 0, ...), invalid lengths leave the owner healthy, several owners.
Full sizes (2^16, 2^20) are checked by round trip plus evaluation of the
coefficient polynomial at sampled roots (X_k = f(w^k)).
"""
import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.codec import encode_fr
from dvgpu.rng import ChaCha20Stream
from oracle import bn254 as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


@pytest.mark.parametrize("seed", [7, 17, 2026])
def test_matches_oracle_across_sizes(backend, seed):
    rng = ChaCha20Stream(seed)
    for n in [1, 2, 8, 16, 128, 256, 1024, 4096]:
        print(f"case: seed={seed} n={n}", flush=True)
        vals = rng.uniform_fr(n)
        vals[0] = o.R - 1
        for inverse in (False, True):
            exp = o.ntt(vals, inverse)
            got = list(vals)
            backend.ntt(got, inverse)
            assert got == exp, (seed, n, inverse)
        rt = list(vals)
        backend.ntt(rt, False)
        backend.ntt(rt, True)
        assert rt == vals


def test_bruteforce_dft_small(backend):
    rng = ChaCha20Stream(3)
    for n in (2, 4, 32):
        vals = rng.uniform_fr(n)
        for inv in (False, True):
            got = list(vals)
            backend.ntt(got, inv)
            assert got == o.dft_bruteforce(vals, inv)


def test_known_values_and_invalid_lengths(backend):
    imp = [1] + [0] * 7
    backend.ntt(imp, False)
    assert imp == [1] * 8
    const = [5] * 8
    backend.ntt(const, False)
    assert const == [40] + [0] * 7
    backend.ntt(const, True)
    assert const == [5] * 8
    for n in (3, 17):
        bad = [7] * n
        with pytest.raises(ValueError):
            backend.ntt(bad, False)
        assert bad == [7] * n
    with pytest.raises(BackendError) as e:
        backend.ntt_encoded(b"\0" * 32 * 3, 3, False)
    assert e.value.status == -1
    # A root of the wrong order is rejected (icicle_msm.cu:783-786).
    with pytest.raises(BackendError) as e:
        backend.ntt_encoded(b"\0" * 64, 2, False, root=encode_fr(o.fr_root_of_unity(1 << 19)))
    assert e.value.status == -1
    healthy = [1] * 16
    backend.ntt(healthy, False)
    assert healthy == [16] + [0] * 15


def test_owners_are_independent():
    survivor = HipBackend()
    vals = [3] * 1024
    survivor.ntt(vals, False)
    for _ in range(3):
        tmp = HipBackend()
        ind = [2] * 256
        tmp.ntt(ind, False)
        assert ind[0] == 512
        tmp.close()
        survivor.ntt(vals, True)
        assert vals == [3] * 1024
        survivor.ntt(vals, False)
    survivor.close()


def _horner(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % o.R
    return acc


@pytest.mark.parametrize("log_n", [14, 16, 20])
def test_full_size_roundtrip_and_sampled_evaluations(backend, log_n):
    n = 1 << log_n
    print(f"case: full size n=2^{log_n}", flush=True)
    rng = ChaCha20Stream(1000 + log_n)
    arr = rng.uniform_fr_array(n)
    enc = arr.tobytes()
    fwd = backend.ntt_encoded(enc, n, False)
    coeffs = [int.from_bytes(enc[32 * i:32 * i + 32], "little") for i in range(n)]
    w = o.fr_root_of_unity(n)
    for k in (0, 1, 2, n // 2, n - 1, 12345 % n):
        got = int.from_bytes(fwd[32 * k:32 * k + 32], "little")
        assert got == _horner(coeffs, pow(w, k, o.R)), k
    back = backend.ntt_encoded(fwd, n, True)
    assert back == enc
