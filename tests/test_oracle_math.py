"""Pins the oracle's deterministic elementary functions (the stand-ins for
libm acos / cos / sin / pow(x, 1.0/3.0) used by isCoplanar and the
impact-zone rotation, dcollid3d.cpp:163-169, 436-444).

They claim to be correctly rounded.  That is checked against values computed
independently here with 70-digit decimal arithmetic (Taylor series for
sin/cos with pi from Machin's formula, Newton on cos for acos, Decimal's own
ln/exp for the power), on random arguments and on every argument where the
host libm disagrees.  glibc is correctly rounded for all but a small
fraction of arguments (its documented worst case is about 0.55 ulp), so the
two agree bit for bit almost everywhere and differ by one ulp elsewhere."""
import math
from decimal import Decimal as D, getcontext
from fractions import Fraction

import numpy as np
import pytest


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


# ---- 70-digit references -------------------------------------------------
getcontext().prec = 80


def _machin_pi():
    def arctan_inv(x):
        x = D(x)
        s, t, n, x2, sg = D(0), 1 / x, 1, x * x, 1
        while True:
            term = t / n
            if term < D(10) ** -85:
                return s
            s += sg * term
            t /= x2
            n += 2
            sg = -sg
    return 4 * (4 * arctan_inv(5) - arctan_inv(239))


PI = _machin_pi()


def _series(r, start, first):
    s, t, n = first, first, start
    while True:
        n += 2
        t = -t * r * r / (n * (n - 1))
        s += t
        if abs(t) < D(10) ** -78:
            return s


def dec_sin(x):
    x = D(x)
    k = int((x / (PI / 2)).to_integral_value())
    r = x - k * (PI / 2)
    c, s = _series(r, 0, D(1)), _series(r, 1, r)
    return [s, c, -s, -c][k % 4]


def dec_cos(x):
    x = D(x)
    k = int((x / (PI / 2)).to_integral_value())
    r = x - k * (PI / 2)
    c, s = _series(r, 0, D(1)), _series(r, 1, r)
    return [c, -s, -c, s][k % 4]


def dec_acos(x):
    y = D(math.acos(x))
    for _ in range(6):
        s = dec_sin(y)
        if s == 0:
            break
        y = y + (dec_cos(y) - D(x)) / s
    return y


THIRD = Fraction(1.0 / 3.0)


def dec_pow_third(x):
    e = D(THIRD.numerator) / D(THIRD.denominator)
    return (e * D(x).ln()).exp()


TRUTH = {"acos": dec_acos, "cos": dec_cos, "sin": dec_sin, "pow_third": dec_pow_third}


def args(name, n, seed):
    rng = np.random.default_rng(seed)
    if name == "acos":
        return np.concatenate([rng.uniform(-1, 1, n), 1 - rng.uniform(0, 1e-6, n // 8),
                               -1 + rng.uniform(0, 1e-6, n // 8)])
    if name in ("cos", "sin"):
        # the cubic's (theta + 2k pi) / 3 lie in [-2.1, 2.1]; zone rotations dt |w|
        return np.concatenate([rng.uniform(-2.2, 2.2, n), rng.uniform(-1e-3, 1e-3, n // 8),
                               rng.uniform(-30, 30, n // 8)])
    return np.concatenate([10.0 ** rng.uniform(-300, 300, n // 2), 10.0 ** rng.uniform(-30, 10, n)])


@pytest.mark.parametrize("name", ["acos", "cos", "sin", "pow_third"])
def test_correctly_rounded_on_random_arguments(O, name):
    x = args(name, 400, seed=11)
    y = O.math_eval(name, x)
    for xi, yi in zip(x, y):
        assert float(TRUTH[name](xi)) == yi, (name, xi)   # float(Decimal) rounds correctly


@pytest.mark.parametrize("name", ["acos", "cos", "sin", "pow_third"])
def test_libm_agrees_except_where_libm_misrounds(O, name):
    x = args(name, 400000, seed=12)
    ours, libm = O.math_eval(name, x), O.math_eval(name, x, libm=True)
    differ = np.nonzero(ours != libm)[0]
    assert len(differ) < 2e-3 * len(x), len(differ)          # agree bit for bit almost always
    ulp = np.abs(ours[differ] - libm[differ]) / np.spacing(np.abs(libm[differ]))
    assert np.all(ulp <= 1.0)
    for i in differ[:25]:                                     # and where not, libm is the one off
        assert float(TRUTH[name](x[i])) == ours[i], (name, x[i])


def test_special_values(O):
    L = O.lib()
    assert L.orc_math_acos(1.0) == 0.0
    assert L.orc_math_acos(-1.0) == math.pi
    assert L.orc_math_acos(0.0) == math.acos(0.0)
    assert math.isnan(L.orc_math_acos(1.0 + 2 ** -52))
    assert math.isnan(L.orc_math_acos(float("nan")))
    assert L.orc_math_cos(0.0) == 1.0 and L.orc_math_sin(0.0) == 0.0
    assert math.copysign(1.0, L.orc_math_sin(-0.0)) == -1.0
    assert L.orc_math_pow_third(0.0) == 0.0
    assert L.orc_math_pow_third(math.inf) == math.inf
    # pow(8, 1.0/3.0) is not 2: the exponent is 1/3 - 2^-54/3
    assert L.orc_math_pow_third(8.0) == 8.0 ** (1.0 / 3.0)
    assert L.orc_math_pow_third(27.0) == 27.0 ** (1.0 / 3.0)


def test_exponent_of_the_reference_cube_root():
    """The reference writes pow(x, 1.0/3.0); the double 1.0/3.0 is below 1/3
    by exactly 2^-54 / 3, the constant the power is corrected by."""
    assert Fraction(1, 3) - THIRD == Fraction(1, 3 * 2 ** 54)
