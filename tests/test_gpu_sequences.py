"""The benchmark's update sequences end to end on the MI355X, each result
checked against the oracle.

* The plain n = 256 sequence (seed 1; 8 updates of k = 4 rewritten
  coefficients, hotspot over a fixed 16-index set or spreading over all 256):
  every update's commitment, evaluations and opening through the per-call
  drop-in equal the oracle's closed forms [f(tau)] g1, NTT(f) and
  [q(tau)] g1, and every proof passes its pairing check.
* The incremental n = 2^16 session (seed 5; 16 updates, k cycling 1, 16,
  256, 4096, hotspot over a fixed 1% subset alternating with spreading):
  every update's proof (C', W', y') equals the trapdoor closed form, the
  resident evaluations equal the C oracle's NTT of the coefficients, the
  proof passes its pairing check, and the plain one-call path gives the same
  bytes.  A proof with a changed opening or value is rejected.
"""
import pytest

from dvgpu import HipBackend
from dvgpu.benchmarks import incremental_updates, small_sequence_gpu
from dvgpu.codec import FR_MODULUS as R, encode_fr, root_of_unity_2_20
from dvgpu.pairing import verify_update_proofs
from dvgpu.serving import PlainSession, SessionSpec, SessionStore, update_batch
from oracle import bn254 as o
from oracle import cref
from tests.test_gpu_serving import Mirror

pytestmark = pytest.mark.gpu
ROOT = encode_fr(root_of_unity_2_20())


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


@pytest.mark.parametrize("workload", ["hotspot", "spreading"])
def test_small_sequence_against_oracle(backend, workload):
    print(f"case: n=256 sequence, {workload}", flush=True)
    run = small_sequence_gpu(backend, workload)
    tau, z, f_final = run["tau"], run["z"], run["f"]
    from dvgpu.workload import small_sequence
    _, coeffs, ups = small_sequence(1, 256, 8, 4, workload)
    f = [int.from_bytes(coeffs[i].tobytes(), "little") for i in range(256)]
    assert len(run["records"]) == 8
    for up, rec in zip(ups, run["records"]):
        for i, v in up:
            f[i] = v
        ft, fz = o.poly_eval(f, tau), o.poly_eval(f, z)
        assert rec["C"] == o.g1_to_bytes(o.g1_mul(ft))
        q = (ft - fz) * pow((tau - z) % R, R - 2, R) % R
        assert rec["W"] == o.g1_to_bytes(o.g1_mul(q))
        assert rec["y"] == fz
        assert rec["evals"] == b"".join(v.to_bytes(32, "little") for v in o.ntt(f))
        assert rec["incremental_ok"] and rec["roundtrip_ok"] and rec["verified"]
    assert f == f_final


def test_incremental_2p16_session(backend):
    spec = SessionSpec(0, 1 << 16, 5)
    store = SessionStore(backend, 1 << 30)
    s = store.admit(spec)
    plain = PlainSession(backend, spec)
    mirror = Mirror(spec)
    ups = incremental_updates(spec.n)
    proofs = []
    for j, (idx, raw, dv) in enumerate(ups):
        print(f"case: incremental update {j}, k={len(idx)}", flush=True)
        p = update_batch(backend, store, [s.update_struct(idx, raw)])[0]
        mirror.apply(idx, dv)
        assert p == mirror.expect(), j
        assert p == plain.apply(backend, idx, raw, dv), j
        ev = store.evaluations(s)
        assert ev == cref.ntt(mirror.coeff_bytes(), spec.n, False, ROOT, 8), j
        assert ev[32 * mirror.t:32 * mirror.t + 32] == p[128:160]
        proofs.append(p)
    assert verify_update_proofs(backend, proofs, [s.divisor] * len(proofs)) == [True] * 16
    # a changed opening, a changed value and a changed commitment are each rejected
    g = o.g1_to_bytes(o.g1_mul(7))
    bad_w = proofs[-1][:64] + g + proofs[-1][128:]
    y = int.from_bytes(proofs[-1][128:160], "little")
    bad_y = proofs[-1][:128] + ((y + 1) % R).to_bytes(32, "little")
    bad_c = g + proofs[-1][64:]
    assert verify_update_proofs(backend, [bad_w, bad_y, bad_c, proofs[0]], [s.divisor] * 4) == \
        [False, False, False, True]
    store.close()


def test_kzg_verify_input_checks(backend):
    """Non-canonical values and off-curve points fail the call (-1) and the
    owner stays usable; identities follow the group law."""
    from dvgpu import BackendError
    tau, z = 1234567, 89
    f_tau, f_z = 1000, 10
    q = (f_tau - f_z) * pow(tau - z, R - 2, R) % R
    d = o.g2_to_bytes(o.g2_mul(tau - z))
    good = o.g1_to_bytes(o.g1_mul(f_tau)) + o.g1_to_bytes(o.g1_mul(q)) + f_z.to_bytes(32, "little")
    assert verify_update_proofs(backend, [good], [d]) == [True]
    with pytest.raises(BackendError):
        verify_update_proofs(backend, [good[:128] + R.to_bytes(32, "little")], [d])
    off = (1).to_bytes(32, "little") + (1).to_bytes(32, "little")
    with pytest.raises(BackendError):
        verify_update_proofs(backend, [off + good[64:]], [d])
    with pytest.raises(BackendError):
        verify_update_proofs(backend, [good], [bytes(64) + (5).to_bytes(32, "little") + bytes(32)])
    # C = [y] g1 and W = O: e(O, g2) == e(O, D) holds for any divisor
    trivial = o.g1_to_bytes(o.g1_mul(f_z)) + bytes(64) + f_z.to_bytes(32, "little")
    assert verify_update_proofs(backend, [trivial, good], [d, d]) == [True, True]
    assert verify_update_proofs(backend, [], []) == []


def test_kzg_verify_many_divisors_both_paths(backend):
    """4200 proofs over 4200 distinct divisors [tau - z_i] g2: the one-wave
    path runs them in two chunks of the 4096-point line cache (the second
    after the cache starts over); 16 385 items (above the one-wave limit)
    take the one-lane path.  Every fifth item carries a changed value and is
    rejected, the others hold (the points come from the fixed-base kernel,
    checked against the oracle elsewhere; the decisions are fixed by the
    algebra)."""
    from dvgpu.pairing import kzg_verify_encoded
    tau, n = 0x1234567890ABCDEF, 4200
    z = [(7 * i + 3) % R for i in range(n)]
    a = [(11 * i + 5) % R for i in range(n)]
    y = [(13 * i + 1) % R for i in range(n)]
    w = [(a[i] - y[i]) * pow(tau - z[i], R - 2, R) % R for i in range(n)]
    enc = lambda vals: b"".join(encode_fr(v) for v in vals)  # noqa: E731
    d = b"".join(backend.fixed_base_encoded(enc([(tau - zz) % R for zz in z[c:c + 4096]]),
                                            len(z[c:c + 4096]), True) for c in range(0, n, 4096))
    c = b"".join(backend.fixed_base_encoded(enc(a[k:k + 4096]), len(a[k:k + 4096]), False) for k in range(0, n, 4096))
    wp = b"".join(backend.fixed_base_encoded(enc(w[k:k + 4096]), len(w[k:k + 4096]), False) for k in range(0, n, 4096))
    vals = enc([(y[i] + (1 if i % 5 == 0 else 0)) % R for i in range(n)])
    want = bytes(0 if i % 5 == 0 else 1 for i in range(n))
    print("case: 4200 distinct divisors, one-wave path", flush=True)
    res, _ = kzg_verify_encoded(backend, c, vals, wp, d, n)
    assert res == want
    print("case: 16385 items, one-lane path", flush=True)
    m = 16385
    rep = lambda buf, size: (buf * (m // n + 1))[:m * size]  # noqa: E731
    res, _ = kzg_verify_encoded(backend, rep(c, 64), rep(vals, 32), rep(wp, 64), rep(d, 128), m)
    assert res == (want * (m // n + 1))[:m]
