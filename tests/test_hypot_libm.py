"""The device path's edge length (csrc/hypot_libm.h, built for the host here) equals the C
library's hypot, which the reference's traction quadrature calls (bc.cpp:231).  That hypot
is not correctly rounded, so the test also shows that a correctly rounded one would differ:
the equality is a property of the algorithm, not of rounding."""
import ctypes
import ctypes.util
import os
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "cpp", "_build", "libhypot_libm.so")


@pytest.fixture(scope="module")
def dev_hypot():
    if not os.path.exists(LIB):
        pytest.skip("tests/cpp/_build/libhypot_libm.so is not built")
    L = ctypes.CDLL(LIB)
    P = ctypes.POINTER(ctypes.c_double)
    L.hypot_batch.argtypes = [ctypes.c_int, P, P, P]

    def f(x, y):
        x, y = np.ascontiguousarray(x, float), np.ascontiguousarray(y, float)
        out = np.empty_like(x)
        L.hypot_batch(x.size, x.ctypes.data_as(P), y.ctypes.data_as(P), out.ctypes.data_as(P))
        return out
    return f


@pytest.fixture(scope="module")
def libm():
    m = ctypes.CDLL(ctypes.util.find_library("m"))
    m.hypot.restype = ctypes.c_double
    m.hypot.argtypes = [ctypes.c_double, ctypes.c_double]
    return np.vectorize(m.hypot, otypes=[float])


def correctly_rounded(h, x, y):
    s = Fraction(x) ** 2 + Fraction(y) ** 2
    lo, hi = np.nextafter(h, 0.0), np.nextafter(h, np.inf)
    return (Fraction(h) + Fraction(lo)) ** 2 / 4 < s < (Fraction(h) + Fraction(hi)) ** 2 / 4


def edge_like(rng, n):
    """Coordinate differences of mesh edges: points in [-10, 10]^2, lengths 1e-4 .. 1."""
    xa, ya = rng.uniform(-10, 10, n), rng.uniform(-10, 10, n)
    h = 10.0 ** rng.uniform(-4, 0, n)
    ang = rng.uniform(0, 2 * np.pi, n)
    return (xa + h * np.cos(ang)) - xa, (ya + h * np.sin(ang)) - ya


def test_equals_libm_on_edges(dev_hypot, libm):
    rng = np.random.default_rng(7)
    dx, dy = edge_like(rng, 200000)
    h = dev_hypot(dx, dy)
    assert np.array_equal(h, libm(dx, dy))
    # not a rounding property: some of these are not the correctly rounded value
    wrong = [i for i in range(3000) if not correctly_rounded(h[i], dx[i], dy[i])]
    assert wrong


def test_equals_libm_wide_range(dev_hypot, libm):
    rng = np.random.default_rng(11)
    n = 50000
    x = rng.uniform(1, 2, n) * 2.0 ** rng.integers(-1070, 1020, n)
    y = x * rng.uniform(0, 1.5, n) * 2.0 ** rng.integers(-60, 3, n)
    assert np.array_equal(dev_hypot(x, y), libm(x, y))


def test_special_values(dev_hypot, libm):
    x = np.array([0.0, -3.0, np.inf, 1.0, np.nan, 0.0, 5e-324, 1e308, 1e-300, 3e200])
    y = np.array([0.0, 4.0, 1.0, -np.inf, 1.0, -2.5, 5e-324, 1e308, 2e-310, 1e-200])
    h = dev_hypot(x, y)
    assert h[0] == 0.0 and h[1] == 5.0 and h[2] == np.inf and h[3] == np.inf
    assert np.isnan(h[4]) and h[5] == 2.5
    finite = np.isfinite(x) & np.isfinite(y)
    assert np.array_equal(h[finite], libm(x[finite], y[finite]))
