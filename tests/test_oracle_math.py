"""Pins the oracle's deterministic elementary functions (the stand-ins for
libm acos / cos / sin / pow(x, 1/3) used by isCoplanar and the impact-zone
rotation) against the platform libm and against exact identities."""
import math

import numpy as np
import pytest


def ulps(a, b):
    a, b = float(a), float(b)
    if a == b:
        return 0
    return abs(a - b) / max(np.spacing(abs(b)), 5e-324)


@pytest.fixture(scope="module")
def L(oracle_lib):
    return oracle_lib.lib()


def test_acos_matches_libm(L):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-1, 1, 20000), [-1.0, -0.5, 0.0, 0.5, 1.0],
                         1 - rng.uniform(0, 1e-6, 100), -1 + rng.uniform(0, 1e-6, 100)])
    worst = max(ulps(L.orc_math_acos(x), math.acos(x)) for x in xs)
    assert worst <= 4, worst
    assert L.orc_math_acos(1.0) == 0.0
    assert math.isnan(L.orc_math_acos(1.0 + 1e-15))
    assert math.isnan(L.orc_math_acos(float("nan")))


def test_cos_sin_match_libm(L):
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-7, 7, 20000), rng.uniform(-1e-3, 1e-3, 1000),
                         [0.0, math.pi / 3, 2 * math.pi / 3, math.pi]])
    wc = max(ulps(L.orc_math_cos(x), math.cos(x)) for x in xs if abs(math.cos(x)) > 1e-3)
    ws = max(ulps(L.orc_math_sin(x), math.sin(x)) for x in xs if abs(math.sin(x)) > 1e-3)
    assert wc <= 4 and ws <= 4, (wc, ws)
    # near zeros of cos/sin compare absolutely (Cody-Waite reduction limit)
    for x in xs:
        assert abs(L.orc_math_cos(x) - math.cos(x)) < 1e-15
        assert abs(L.orc_math_sin(x) - math.sin(x)) < 1e-15


def test_cbrt_is_the_cube_root(L):
    rng = np.random.default_rng(3)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 20000), [0.0, 1.0, 8.0, 27.0, 1e-300]])
    worst = max(ulps(L.orc_math_cbrt(x), np.cbrt(x)) for x in xs)
    assert worst <= 2, worst
    assert L.orc_math_cbrt(8.0) == 2.0 and L.orc_math_cbrt(27.0) == 3.0
    assert L.orc_math_cbrt(0.0) == 0.0


def test_cbrt_vs_reference_pow_one_third(L):
    """The reference writes pow(x, 1.0/3.0); the double 1.0/3.0 is below 1/3 by
    d = 1.85e-17, so pow differs from the cube root by a factor exp(-d ln x).
    Over the magnitudes the coplanarity cubic meets (|R| in 1e-30..1e10) the two
    agree to within that bound plus a few ulps."""
    from fractions import Fraction
    d = float(Fraction(1, 3) - Fraction(1.0 / 3.0))
    assert 1.8e-17 < d < 1.9e-17
    rng = np.random.default_rng(4)
    for x in 10.0 ** rng.uniform(-30, 10, 5000):
        a, b = L.orc_math_cbrt(x), x ** (1.0 / 3.0)
        bound = abs(b) * (d * abs(math.log(x)) + 6e-16)
        assert abs(a - b) <= bound, (x, a, b)
