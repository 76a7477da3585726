"""Polynomial tools with batched device products (dynaverse_poly.h) against
the quadratic oracles of the reference's own tests
(crates/protocol/src/domain/fast_tests.rs) on the same shapes."""
import random

import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.poly import SubproductTree, fft_multiply, newton_divide_exact
from oracle import bn254 as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


@pytest.mark.parametrize("size", [0, 1, 2, 3, 16, 17, 33, 65])
def test_tree_product_evaluation_and_interpolation(backend, size):
    """fast_tests.rs:24-42: points 3i + 2, degrees 0, size, 2 size + 3."""
    rng = random.Random(15201 + size)
    points = [3 * i + 2 for i in range(size)]
    with SubproductTree(backend, points) as tree:
        van = [1]
        for x in points:
            van = o.poly_mul(van, [(-x) % o.R, 1])
        assert tree.vanishing() == van
        for degree in (0, size, 2 * size + 3):
            poly = [rng.randrange(o.R) for _ in range(degree + 1)]
            values = [o.poly_eval(poly, x) for x in points]
            assert tree.multipoint_eval(poly) == values
            assert tree.interpolate(values) == o.lagrange_interpolate(points, values)
        assert tree.interpolate([0] * size) == [0]
        with pytest.raises(BackendError):
            tree.interpolate([0] * (size + 1))
        der = o.poly_trim([i * c for i, c in enumerate(van)][1:])
        assert tree.inverse_derivatives() == [pow(o.poly_eval(der, x), o.R - 2, o.R) for x in points]


def test_duplicate_points_rejected(backend):
    with pytest.raises(BackendError) as e:
        SubproductTree(backend, [1, 1])
    assert e.value.status == -1


@pytest.mark.parametrize("size", [0, 1, 2, 3, 17, 32, 65])
def test_cauchy_sums(backend, size):
    """fast_tests.rs:44-62: points 2(size - i) + 1."""
    rng = random.Random(15202 + size)
    points = [2 * (size - i) + 1 for i in range(size)]
    weights = [rng.randrange(o.R) for _ in range(size)]
    with SubproductTree(backend, points) as tree:
        assert tree.cauchy_sums(weights) == o.cauchy_sums_direct(points, weights)
        assert tree.cauchy_sums([0] * size) == [0] * size
        with pytest.raises(BackendError):
            tree.cauchy_sums([0] * (size + 1))


@pytest.mark.parametrize("degree", [0, 1, 3, 17, 33, 65])
def test_newton_division(backend, degree):
    """fast_tests.rs:64-86: dense quotients, every corrupted low coefficient rejected."""
    rng = random.Random(15203 + degree)
    den = [rng.randrange(o.R) for _ in range(degree)] + [3]
    for length in (1, 3, 17, 67):
        q = [rng.randrange(o.R) for _ in range(length)]
        num = o.poly_mul(den, q)
        assert newton_divide_exact(backend, num, den) == o.poly_trim(q) == o.poly_divide_exact(num, den)
        for index in range(min(degree, 4)):
            bad = list(num)
            bad[index] = (bad[index] + 1) % o.R
            with pytest.raises(BackendError):
                newton_divide_exact(backend, bad, den)
    assert newton_divide_exact(backend, [], [1]) == [0]
    with pytest.raises(BackendError):
        newton_divide_exact(backend, [], [])
    with pytest.raises(BackendError):
        newton_divide_exact(backend, [1], [0, 1])


def test_fft_multiply(backend):
    """fast_tests.rs:88-96 and random shapes."""
    assert fft_multiply(backend, [1] * 17, [2] * 17) == o.poly_mul([1] * 17, [2] * 17)
    rng = random.Random(44)
    for la, lb in ((1, 1), (2, 3), (100, 37), (1000, 1)):
        a = [rng.randrange(o.R) for _ in range(la)]
        b = [rng.randrange(o.R) for _ in range(lb)]
        assert fft_multiply(backend, a, b) == o.poly_mul(a, b)
    assert fft_multiply(backend, [], [5]) == [0]


def test_large_tree_batches_levels(backend):
    """4096 points: values against Horner at sampled points, interpolation
    round trip, one device round per tree level for the products."""
    rng = random.Random(45)
    n = 4096
    points = rng.sample(range(1, 1 << 40), n)
    with SubproductTree(backend, points) as tree:
        build = tree.stats()
        assert build["products"] >= n - 1
        poly = [rng.randrange(o.R) for _ in range(n)]
        vals = tree.multipoint_eval(poly)
        for i in (0, 1, 777, n - 1):
            assert vals[i] == o.poly_eval(poly, points[i])
        assert tree.interpolate(vals) == o.poly_trim(poly)
        inter = tree.stats()
        assert inter["rounds"] <= 12                       # at most one device round per level
    print("tree build", build, "interpolate", inter, flush=True)


def test_mid_size_tree_device_rounds_checked_everywhere(backend):
    """400 points, a tree kept in HBM (nodes up to 256 leaves, the last
    nodes of every level partly filled): the vanishing polynomial, the
    values at every point, the interpolation, the inverse derivatives and the
    Cauchy sums all against the oracle's direct forms; Newton division with a
    degree-300 divisor through the batched device products."""
    rng = random.Random(46)
    n = 400
    points = rng.sample(range(1, 1 << 40), n)
    with SubproductTree(backend, points) as tree:
        assert tree.stats()["device_products"] >= 1        # device rounds, not the host schoolbook
        van = [1]
        for x in points:
            van = o.poly_mul(van, [(-x) % o.R, 1])
        assert tree.vanishing() == van
        poly = [rng.randrange(o.R) for _ in range(n)]
        vals = tree.multipoint_eval(poly)
        assert vals == [o.poly_eval(poly, x) for x in points]
        assert tree.interpolate(vals) == o.poly_trim(poly)
        der = o.poly_trim([i * c for i, c in enumerate(van)][1:])
        assert tree.inverse_derivatives() == [pow(o.poly_eval(der, x), o.R - 2, o.R) for x in points]
        weights = [rng.randrange(o.R) for _ in range(n)]
        assert tree.cauchy_sums(weights) == o.cauchy_sums_direct(points, weights)
    den = [rng.randrange(o.R) for _ in range(300)] + [7]
    q = [rng.randrange(o.R) for _ in range(257)]
    num = o.poly_mul(den, q)
    assert newton_divide_exact(backend, num, den) == o.poly_trim(q)
    bad = list(num)
    bad[5] = (bad[5] + 1) % o.R
    with pytest.raises(BackendError):
        newton_divide_exact(backend, bad, den)


@pytest.mark.parametrize("n", [64, 513])
def test_device_tree_boundary_shapes(backend, n):
    """Trees in HBM where they start (64 points, a power of two) and with a
    one-point node at the top split (513 = 512 + 1): every output against the
    oracle's direct forms, and a polynomial longer than the point count
    (reduced modulo the vanishing polynomial before the remainder tree)."""
    rng = random.Random(47 + n)
    points = rng.sample(range(1, 1 << 40), n)
    with SubproductTree(backend, points) as tree:
        assert tree.stats()["device_products"] >= 1
        van = [1]
        for x in points:
            van = o.poly_mul(van, [(-x) % o.R, 1])
        assert tree.vanishing() == van
        poly = [rng.randrange(o.R) for _ in range(n)]
        vals = tree.multipoint_eval(poly)
        assert vals == [o.poly_eval(poly, x) for x in points]
        assert tree.interpolate(vals) == o.poly_trim(poly)
        longer = [rng.randrange(o.R) for _ in range(2 * n + 3)]
        assert tree.multipoint_eval(longer) == [o.poly_eval(longer, x) for x in points]
        der = o.poly_trim([i * c for i, c in enumerate(van)][1:])
        assert tree.inverse_derivatives() == [pow(o.poly_eval(der, x), o.R - 2, o.R) for x in points]
        weights = [rng.randrange(o.R) for _ in range(n)]
        assert tree.cauchy_sums(weights) == o.cauchy_sums_direct(points, weights)
        assert tree.interpolate([0] * n) == [0]


def test_device_tree_keeps_its_root(backend):
    """A call naming another root of order 2^20 gets the same values: the
    tree's transformed rows were made with the root of its build and its
    rounds keep to that root; the owner's later products agree too."""
    import ctypes

    from dvgpu import _lib
    from dvgpu.poly import _buf, _dec, _enc
    rng = random.Random(49)
    n = 100
    points = rng.sample(range(1, 1 << 40), n)
    poly = [rng.randrange(o.R) for _ in range(n)]
    expect = [o.poly_eval(poly, x) for x in points]
    other = _buf(o.fr_to_bytes(pow(o.ROOT_2_20, 3, o.R)))
    with SubproductTree(backend, points) as tree:
        out = ctypes.create_string_buffer(32 * n)
        err = ctypes.create_string_buffer(256)
        st = _lib.DvgpuPolyStats()
        assert backend.lib.dvgpu_poly_tree_eval(backend.handle, other, tree.handle, _buf(_enc(poly)), n, out,
                                                ctypes.byref(st), err, 256) == 0, err.value
        assert _dec(out.raw, n) == expect
        assert tree.multipoint_eval(poly) == expect
        assert tree.interpolate(expect) == o.poly_trim(poly)
    a = [rng.randrange(o.R) for _ in range(70)]
    assert fft_multiply(backend, a, poly) == o.poly_mul(a, poly)


def test_device_tree_at_the_point_limit(backend):
    """2^19 points, the largest tree (the entry point refuses one more):
    values of a degree-(2^19 - 1) polynomial spot-checked by Horner at both
    ends and inside, and the interpolation of all 2^19 values giving the
    polynomial back exactly (a size-independent round trip)."""
    rng = random.Random(51)
    n = 1 << 19
    points = rng.sample(range(1, 1 << 60), n)
    poly = [rng.randrange(o.R) for _ in range(n)]
    with SubproductTree(backend, points) as tree:
        vals = tree.multipoint_eval(poly)
        for i in (0, 1, 4097, n // 2, n - 2, n - 1):
            assert vals[i] == o.poly_eval(poly, points[i]), i
        assert tree.interpolate(vals) == o.poly_trim(poly)
    with pytest.raises(BackendError):
        SubproductTree(backend, points + [(1 << 60) + 5])
