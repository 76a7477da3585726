"""Host logic of the serving form (no GPU): range allocation, update merging,
seeded request streams and session derivation."""
import numpy as np

from dvgpu.codec import FR_MODULUS as R, root_of_unity_2_20
from dvgpu.serving import (RangeAllocator, SessionSpec, merge_update, percentile, request_stream,
                           session_bytes, session_pool, session_scalars)


def test_range_allocator_first_fit_and_merge():
    a = RangeAllocator(100)
    x, y, z = a.alloc(30), a.alloc(30), a.alloc(30)
    assert (x, y, z) == (0, 30, 60) and a.alloc(20) is None
    a.release(y, 30)
    assert a.alloc(40) is None and a.alloc(30) == 30
    a.release(x, 30)
    a.release(30, 30)
    a.release(z, 30)
    assert a.free == [(0, 100)] and a.free_units == 100


def test_merge_update_sums_repeats():
    idx, raw, vals = merge_update([5, 2, 5, 9, 2], [1, 2, 3, R - 1, R - 2])
    assert idx.tolist() == [2, 5, 9]
    assert vals == [0, 4, R - 1]
    assert raw == b"".join(v.to_bytes(32, "little") for v in vals)


def test_session_scalars_relations():
    spec = SessionSpec(0, 16, 123)
    tau, t, z, coeffs = spec.derive()
    w16 = pow(root_of_unity_2_20(), 1 << 16, R)
    assert pow(w16, 16, R) == 1 and z == pow(w16, t, R)
    a, u, l = session_scalars(tau, z, 16)
    for i in range(16):
        ai = int.from_bytes(a[i].tobytes(), "little")
        ui = int.from_bytes(u[i].tobytes(), "little")
        li = int.from_bytes(l[i].tobytes(), "little")
        assert ai == pow(tau, i, R) and li == pow(z, i, R)
        assert ui * (tau - z) % R == (ai - li) % R        # U_i = (tau^i - z^i)/(tau - z)
    # Opening relation for f: sum f_i U_i = (f(tau) - f(z))/(tau - z).
    f = [int.from_bytes(coeffs[i].tobytes(), "little") for i in range(16)]
    ft = sum(c * pow(tau, i, R) for i, c in enumerate(f)) % R
    fz = sum(c * pow(z, i, R) for i, c in enumerate(f)) % R
    q = sum(c * int.from_bytes(u[i].tobytes(), "little") for i, c in enumerate(f)) % R
    assert q * (tau - z) % R == (ft - fz) % R


def test_request_stream_deterministic_and_mixed():
    specs = session_pool(4)
    a = request_stream(specs, 400, 1000.0)
    b = request_stream(specs, 400, 1000.0)
    assert [(x[0], x[1], list(x[2]), list(x[3])) for x in a] == [(x[0], x[1], list(x[2]), list(x[3])) for x in b]
    sizes = {s.sid: s.n for s in specs}
    counts = {n: 0 for n in (256, 1024, 4096, 65536)}
    for t, sid, idx, ds in a:
        counts[sizes[sid]] += 1
        assert 1 <= len(idx) <= 64 and all(0 <= i < sizes[sid] for i in idx)
        assert all(0 <= d < R for d in ds)
    assert counts[256] > counts[1024] > counts[4096] > counts[65536] > 0
    gaps = np.diff([0.0] + [x[0] for x in a])
    assert abs(gaps.mean() - 1e-3) < 2.5e-4                # Poisson at 1000 QPS


def test_session_bytes_and_percentile():
    assert session_bytes(256) == 64 * 32 * 514 + 32 * 769
    assert percentile([1, 2, 3, 4, 5], 50) == 3 and percentile([1, 2, 3, 4, 5], 100) == 5


def test_coset_point_constant_is_off_subgroup():
    from dvgpu.benchmarks import COSET_POINT
    from oracle import bn254 as o
    from tests.g2corpus import twist_cofactor_point
    T = twist_cofactor_point()
    assert o.g2_to_bytes(T) == COSET_POINT
    assert o.on_curve(T, o.FQ2) and o.mul(T, o.R, o.FQ2) is not None
