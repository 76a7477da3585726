"""CPU pins of the polynomial oracle forms (oracle/bn254.py) the GPU tests of
the polynomial tools compare against: the same quadratic algorithms the
reference's fast_tests.rs uses as its oracles, checked against what the
mathematics fixes (products evaluate as products, quotients multiply back,
interpolants hit their values, Cauchy sums against a direct definition on
points with known differences)."""
import random

import pytest

from oracle import bn254 as o


def test_products_and_division():
    rng = random.Random(41)
    for la, lb in ((1, 1), (3, 5), (17, 17), (1, 40)):
        a = [rng.randrange(o.R) for _ in range(la)]
        b = [rng.randrange(o.R) for _ in range(lb)]
        p = o.poly_mul(a, b)
        for x in (0, 1, 12345, o.R - 1):
            assert o.poly_eval(p, x) == o.poly_eval(a, x) * o.poly_eval(b, x) % o.R
        if b[-1]:
            assert o.poly_divide_exact(p, b) == o.poly_trim(a)
    assert o.poly_mul([], [1, 2]) == [0]
    with pytest.raises(ValueError):
        o.poly_divide_exact([1], [0, 1])
    with pytest.raises(ValueError):
        o.poly_divide_exact([], [])
    assert o.poly_divide_exact([], [1]) == [0]


def test_interpolation_and_cauchy_sums():
    rng = random.Random(42)
    pts = [3 * i + 2 for i in range(9)]
    vals = [rng.randrange(o.R) for _ in pts]
    f = o.lagrange_interpolate(pts, vals)
    assert len(f) <= len(pts) and [o.poly_eval(f, x) for x in pts] == vals
    # points 0, 1, 2: sums of w_j/(x_i - x_j) written out
    w = [5, 7, 11]
    inv = lambda v: pow(v % o.R, o.R - 2, o.R)
    exp = [(7 * inv(-1) + 11 * inv(-2)) % o.R, (5 * inv(1) + 11 * inv(-1)) % o.R, (5 * inv(2) + 7 * inv(1)) % o.R]
    assert o.cauchy_sums_direct([0, 1, 2], w) == exp
