"""Device FP64 rounding as compiled for the solver: sqrt and division must
be correctly rounded (bitwise equal to the host's IEEE results), and the
solver's constant-divisor quotient (div_const) must equal the IEEE quotient:
random numerators over a wide exponent range, numerators whose quotients
lie next to rounding midpoints, signed zeros, subnormal numerators,
numerators and quotients next to the normal boundary and next to overflow,
infinities and NaN, for the solver's divisors
(dx, dy, dx^2, dy^2, Vm, 2A, 3) and for divisors with extreme
significands."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
LIB = os.path.join(ROOT, "tests", "hip", "librounding_probe.so")
_dp = C.POINTER(C.c_double)


def lib():
    L = C.CDLL(LIB)
    L.probe_sqrt.argtypes = [_dp, _dp, C.c_int]
    L.probe_div.argtypes = [_dp, _dp, _dp, C.c_int]
    L.probe_div_const.argtypes = [_dp, C.c_double, C.c_double, _dp, C.c_int]
    return L


def p(a):
    return a.ctypes.data_as(_dp)


def random_doubles(rng, n, lo=-300, hi=300):
    m = rng.random(n) + 1.0
    e = rng.integers(lo, hi, n)
    return np.ldexp(m, e)


def test_sqrt_correctly_rounded():
    rng = np.random.default_rng(7)
    x = np.concatenate([random_doubles(rng, 1 << 22), rng.random(1 << 20) * 1e-20, rng.random(1 << 20),
                        np.array([0.0, 1.0, 2.0, 4.0, 1e-310, 5e-324, 1.7976931348623157e308])])
    x = np.ascontiguousarray(x)
    y = np.empty_like(x)
    assert lib().probe_sqrt(p(x), p(y), x.size) == 0
    ref = np.sqrt(x)
    bad = np.flatnonzero(y != ref)
    assert bad.size == 0, (bad.size, x[bad[:5]], y[bad[:5]], ref[bad[:5]])


def test_division_correctly_rounded():
    rng = np.random.default_rng(8)
    a = np.ascontiguousarray(random_doubles(rng, 1 << 22) * rng.choice([-1, 1], 1 << 22))
    b = np.ascontiguousarray(random_doubles(rng, 1 << 22))
    q = np.empty_like(a)
    assert lib().probe_div(p(a), p(b), p(q), a.size) == 0
    assert np.array_equal(q, a / b)


@pytest.mark.parametrize("d", [2.79e-9, 2.79e-9 * 2.79e-9, 1e-9, 1e-9 * 1e-9, 7.1e-6, 3.0, 2.0 * 13021.0,
                               2.0 * 1.5e5, 7.7841e-18, 1.2345678901234567, float.fromhex("0x1.fffffffffffffp0"),
                               float.fromhex("0x1.0000000000001p0"), float.fromhex("0x1.fffffffffffffp-40")])
def test_constant_divisor_quotient(d):
    rng = np.random.default_rng(9)
    n = 1 << 21
    m = 1 << 16
    sign = rng.choice([-1, 1], m)
    tiny = np.finfo(np.float64).smallest_subnormal
    with np.errstate(over="ignore"):
        a = np.concatenate([random_doubles(rng, n, -250, 250) * rng.choice([-1, 1], n),
                            rng.normal(size=n), rng.normal(size=n) * 1e-20,
                            # quotients next to rounding midpoints: a = (m + 1/2 ulp) * d rounded
                            (random_doubles(rng, n, -40, 40) * d),
                            np.nextafter(random_doubles(rng, n, -40, 40) * d, np.inf),
                            # subnormal numerators, numerators and quotients next to the
                            # normal boundary and next to overflow
                            sign * rng.integers(1, 1 << 52, m) * tiny,
                            sign * random_doubles(rng, m, -1075, -940),
                            sign * random_doubles(rng, m, -1022, -1020),
                            sign * random_doubles(rng, m, 990, 1023),
                            sign * random_doubles(rng, m, -960, -900) * d,
                            sign * np.nextafter(np.ldexp(1.0, -960) * d, 0.0) * rng.random(m),
                            sign * np.minimum(random_doubles(rng, m, 1015, 1024) * d, np.finfo(np.float64).max),
                            np.array([0.0, -0.0, 1.0, -1.0, d, -d, tiny, -tiny, np.inf, -np.inf, np.nan,
                                      np.finfo(np.float64).max, -np.finfo(np.float64).max,
                                      np.finfo(np.float64).tiny, np.ldexp(1.0, -960), np.ldexp(1.0, 1000)])])
    a = np.ascontiguousarray(a)
    q = np.empty_like(a)
    assert lib().probe_div_const(p(a), d, 1.0 / d, p(q), a.size) == 0
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        ref = a / d
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(q), nan)
    bad = np.flatnonzero((q != ref) & ~nan)
    assert bad.size == 0, (bad.size, a[bad[:5]])
    assert np.array_equal(np.signbit(q[~nan]), np.signbit(ref[~nan]))
