"""Opening-cache precompute on the MI355X (dynaverse_kzg.h) against the
oracle's restatement of crates/protocol/src/kzg.rs:217-275 and against the
dense openings of the trapdoor (the reference's ``assert_dense_oracle``,
kzg.rs:395-404), with the retained-basis rules of kzg.rs:277-300 and the
reference's test shapes (kzg.rs:366-520)."""
import random

import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.kzg import KzgSrs
from oracle import bn254 as o
from oracle import cref

pytestmark = pytest.mark.gpu

ROOT_2_20 = o.fr_to_bytes(o.fr_root_of_unity(1 << 20))


def coefficients(length, offset):
    return [i * i + 3 * i + offset for i in range(length)]


def dense(tau, coeffs, n, offset=None, indices=None):
    w = o.fr_root_of_unity(n)
    g = 1 if offset is None else offset
    idx = range(n) if indices is None else indices
    return [o.kzg_opening_closed_form(tau, coeffs, g * pow(w, i, o.R) % o.R) for i in idx]


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


@pytest.fixture(scope="module")
def srs16(backend):
    pts = o.kzg_srs(7, 16)
    z = KzgSrs(backend, pts)
    yield z, pts
    z.close()


def test_same_degree_reuses_basis(srs16):
    """kzg.rs:406-420."""
    z, pts = srs16
    first = coefficients(8, 7)
    assert z.precompute_single_openings(first, 8) == dense(7, first, 8)
    deg, builds = z.retained_basis()
    assert deg == 7
    second = coefficients(8, 19)
    assert z.precompute_single_openings(second, 8) == dense(7, second, 8)
    assert z.retained_basis() == (7, builds)
    assert z.precompute_single_openings(first, 8) == dense(7, first, 8)   # basis unchanged by use


def test_degree_switch_rebuilds_one_entry(srs16):
    """kzg.rs:422-437: lengths 8, 7, 4, 8 over a 16-point domain."""
    z, pts = srs16
    prev, builds = z.retained_basis()
    for length in (8, 7, 4, 8):
        c = coefficients(length, 11)
        assert z.precompute_single_openings(c, 16) == o.kzg_openings(pts, c, 16) == dense(7, c, 16)
        deg, now = z.retained_basis()
        assert deg == length - 1 and now == builds + (deg != prev)
        prev, builds = deg, now


def test_coset_domain_keeps_the_basis(srs16):
    """kzg.rs:491-503: a 16-point coset by 5 reuses the degree-6 basis."""
    z, pts = srs16
    c = coefficients(7, 13)
    assert z.precompute_single_openings(c, 8) == dense(7, c, 8)
    before = z.retained_basis()
    assert z.precompute_single_openings(c, 16, offset=5) == dense(7, c, 16, offset=5)
    assert z.retained_basis() == before


def test_constants_and_invalid_inputs_do_not_touch_the_basis(backend):
    """kzg.rs:466-488."""
    pts = o.kzg_srs(7, 8)
    with KzgSrs(backend, pts) as z:
        for poly in ([], [3]):
            assert z.precompute_single_openings(poly, 8) == [None] * 8
        for bad, n in ((coefficients(9, 3), 8), (coefficients(10, 3), 16)):
            with pytest.raises(BackendError) as e:
                z.precompute_single_openings(bad, n)
            assert e.value.status == -1
        assert z.retained_basis() == (0, 0)
        c = coefficients(8, 9)
        assert z.precompute_single_openings(c, 8) == dense(7, c, 8)
        first = z.retained_basis()
        z.precompute_single_openings([], 8)
        with pytest.raises(BackendError):
            z.precompute_single_openings(coefficients(9, 3), 8)
        with pytest.raises(BackendError):
            z.precompute_single_openings(c, 8, offset=0)
        with pytest.raises(BackendError):
            z.precompute_bytes(o.R.to_bytes(32, "little") * 8, 8, 1, 8)
        assert z.retained_basis() == first
        assert z.precompute_single_openings(c, 8) == dense(7, c, 8)


def test_distinct_srs_objects_do_not_share(backend):
    """kzg.rs:439-464: another trapdoor gives another basis and openings."""
    c = coefficients(8, 5)
    with KzgSrs(backend, o.kzg_srs(7, 16)) as a, KzgSrs(backend, o.kzg_srs(11, 16)) as b:
        ra = a.precompute_single_openings(c, 8)
        assert b.retained_basis() == (0, 0)
        rb = b.precompute_single_openings(c, 8)
        assert ra == dense(7, c, 8) and rb == dense(11, c, 8) and ra != rb


@pytest.mark.parametrize("rows", [3, 64])
def test_batched_rows_equal_single_calls(backend, rows):
    """Several caches of one degree in one call equal each computed alone
    (64 rows take the unsplit transform path)."""
    rng = random.Random(rows)
    tau = 31337
    pts = o.kzg_srs(tau, 32)
    n = 32 if rows == 3 else 8
    length = n if rows == 3 else 8
    polys = [[rng.randrange(o.R) for _ in range(length)] for _ in range(rows)]
    with KzgSrs(backend, pts) as z:
        batch = z.precompute_batch(polys, n)
        for r in (0, rows // 2, rows - 1):
            assert batch[r] == z.precompute_single_openings(polys[r], n)
            assert batch[r] == o.kzg_openings(pts, polys[r], n)


@pytest.mark.parametrize("log_n,length", [(0, 1), (1, 2), (2, 3), (5, 17), (6, 64), (8, 200), (10, 1024),
                                          (11, 1500), (12, 4096)])
def test_sizes_against_the_c_oracle(backend, log_n, length):
    """Every transform length from 1 to 2^13 (both stage splits) bit-exact."""
    rng = random.Random(log_n * 1000 + length)
    n = 1 << log_n
    tau = rng.randrange(2, o.R)
    srs = cref.fixed_base_g1(b"".join(o.fr_to_bytes(pow(tau, i, o.R)) for i in range(length)), length)
    c = [rng.randrange(o.R) for _ in range(length)]
    cb = b"".join(o.fr_to_bytes(v) for v in c)
    with KzgSrs(backend, encoded=srs) as z:
        got = z.precompute_bytes(cb, length, 1, n)
    assert got == cref.kzg_openings(srs, length, cb, length, n, ROOT_2_20, threads=16)


def test_full_size_caches_sampled_against_dense_openings(backend):
    """Three caches of degree 2^16 - 1 over 2^16 points (the anchor-opening
    workload, dynaverse.rs:303-314) in one call; sampled positions against
    the trapdoor's dense openings."""
    n = 1 << 16
    rng = random.Random(99)
    tau = rng.randrange(2, o.R)
    powers = [pow(tau, i, o.R) for i in range(n)]
    srs = b"".join(backend.fixed_base_encoded(b"".join(o.fr_to_bytes(v) for v in powers[c:c + 4096]),
                                              len(powers[c:c + 4096]), False) for c in range(0, n, 4096))
    polys = [[rng.randrange(o.R) for _ in range(n)] for _ in range(3)]
    with KzgSrs(backend, encoded=srs) as z:
        raw = z.precompute_bytes(b"".join(o.fr_to_bytes(v) for p in polys for v in p), n, 3, n)
        t = z.last_timings
    print("timings", t, flush=True)
    idx = [0, 1, 2, 12345, n // 2, n - 1]
    for r, p in enumerate(polys):
        exp = dense(tau, p, n, indices=idx)
        for i, e in zip(idx, exp):
            assert raw[64 * (r * n + i):64 * (r * n + i + 1)] == o.g1_to_bytes(e), (r, i)


def test_largest_convolution(backend):
    """The largest supported shape: 2^19 coefficients over 2^19 points, a
    2^20-point convolution (every twiddle of the 2^20 table in use), sampled
    against the trapdoor; one coefficient more than that is refused."""
    n = 1 << 19
    rng = random.Random(77)
    tau = rng.randrange(2, o.R)
    powers = [1] * (n + 2)
    for i in range(1, n + 2):
        powers[i] = powers[i - 1] * tau % o.R
    srs = b"".join(backend.fixed_base_encoded(b"".join(o.fr_to_bytes(v) for v in powers[c:c + 4096]),
                                              len(powers[c:c + 4096]), False) for c in range(0, n + 2, 4096))
    coeffs = [rng.randrange(o.R) for _ in range(n)]
    with KzgSrs(backend, encoded=srs) as z:
        raw = z.precompute_bytes(b"".join(o.fr_to_bytes(v) for v in coeffs), n, 1, n)
        with pytest.raises(BackendError):
            z.precompute_bytes(bytes(32 * (n + 2)), n + 2, 1, 1 << 20)   # convolution of 2^21 points
    idx = [0, 1, 3, n // 2 + 1, n - 1]
    for i, e in zip(idx, dense(tau, coeffs, n, indices=idx)):
        assert raw[64 * i:64 * i + 64] == o.g1_to_bytes(e), i


def test_simultaneous_callers_get_independent_correct_openings(backend):
    """kzg.rs:505-525: two threads on one SRS with different polynomials
    (the SRS lock serialises the calls; each result is its own)."""
    import threading
    pts = o.kzg_srs(7, 16)
    results, errors = {}, []
    with KzgSrs(backend, pts) as z:
        barrier = threading.Barrier(2)

        def run(offset):
            try:
                c = coefficients(8, offset)
                barrier.wait()
                for _ in range(5):
                    results[offset] = z.precompute_single_openings(c, 8)
            except Exception as e:      # surfaced below
                errors.append(e)
        threads = [threading.Thread(target=run, args=(off,)) for off in (3, 17)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors
        assert z.retained_basis()[0] == 7
    for off in (3, 17):
        assert results[off] == dense(7, coefficients(8, off), 8)
