"""Pins the oracle's floor, ceil, trunc, round, round2even and sign (CPU only)
against the known answers of Boost.SIMD's unit tests and against exact
rational arithmetic (fractions.Fraction) on seeded samples that include every
half-integer tie the formats can hold."""
import math
from fractions import Fraction

import numpy as np
import pytest

import _oracle as O

HALF = {np.float32: np.float32(0.5), np.float64: 0.5}


def _map(dt, name, x):
    mapf = O.map_f32 if dt is np.float32 else O.map_f64
    return mapf(name, np.asarray(x, dtype=dt))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_boost_simd_unit_kats(dt):
    inf, nan, h = dt(np.inf), dt(np.nan), dt(0.5)
    below_half = np.nextafter(h, dt(0))
    # modules/boost/simd/base/unit/arithmetic/simd/round.cpp:49-67
    x = [inf, -inf, nan, 1.4, 1.5, 1.6, 2.5, h, -h, -1, 1, 0,
         np.nextafter(below_half, dt(0)), below_half, np.nextafter(h, dt(1))]
    want = [inf, -inf, nan, 1, 2, 2, 3, 1, -1, -1, 1, 0, 0, 0, 1]
    assert np.array_equal(_map(dt, "round", x), np.array(want, dtype=dt), equal_nan=True)
    # round2even.cpp:42-54
    x = [inf, -inf, nan, 1.4, 1.5, 1.6, 2.5, h, -h, -1, 1, 0]
    want = [inf, -inf, nan, 1, 2, 2, 2, 0, 0, -1, 1, 0]
    assert np.array_equal(_map(dt, "round2even", x), np.array(want, dtype=dt), equal_nan=True)
    # floor.cpp:36-44, ceil.cpp:36-44
    x = [inf, -inf, nan, -1, 1, 0, -1.1, 1.1]
    assert np.array_equal(_map(dt, "floor", x), np.array([inf, -inf, nan, -1, 1, 0, -2, 1], dtype=dt),
                          equal_nan=True)
    assert np.array_equal(_map(dt, "ceil", x), np.array([inf, -inf, nan, -1, 1, 0, -1, 2], dtype=dt),
                          equal_nan=True)
    # trunc.cpp:41-53
    x = [inf, -inf, nan, 1.4, 1.5, 1.6, 2.5, h, -h, -1, 1, 0]
    want = [inf, -inf, nan, 1, 1, 1, 2, 0, 0, -1, 1, 0]
    assert np.array_equal(_map(dt, "trunc", x), np.array(want, dtype=dt), equal_nan=True)
    # modules/boost/simd/base/unit/ieee/simd/sign.cpp (real case):
    # sign(Inf) = 1, sign(Minf) = -1, sign(Nan) = Nan, sign(Mone) = -1, sign(One) = 1, sign(Zero) = 0
    got = _map(dt, "sign", [inf, -inf, nan, -1, 1, 0, -0.0, 1e-30, -3e20])
    assert np.array_equal(got, np.array([1, -1, nan, -1, 1, 0, 0, 1, -1], dtype=dt), equal_nan=True)


def _samples(dt, rng):
    n = 4000
    mant = 23 if dt is np.float32 else 52
    ties = (rng.integers(-(1 << 20), 1 << 20, n) + 0.5).astype(dt)  # exact k + 1/2
    small = rng.uniform(-4, 4, n).astype(dt)
    big = (rng.uniform(-1, 1, n) * 2.0 ** rng.uniform(0, mant + 3, n)).astype(dt)
    near = np.nextafter(ties, dt(np.inf) * np.sign(rng.uniform(-1, 1, n)).astype(dt))
    return np.concatenate([ties, -ties, small, big, near, np.array([0.0, -0.0], dtype=dt)])


def _exact(name, v):
    q = Fraction(float(v))
    if name == "floor":
        r = math.floor(q)
    elif name == "ceil":
        r = math.ceil(q)
    elif name == "trunc":
        r = math.trunc(q)
    elif name == "round":  # half away from zero
        r = math.floor(abs(q) + Fraction(1, 2)) * (1 if q >= 0 else -1)
    else:  # round2even: Python's round() on a Fraction is half to even
        r = round(q)
    return math.copysign(float(r), float(v)) if r == 0 else float(r)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("name", ["floor", "ceil", "trunc", "round", "round2even"])
def test_rounding_against_exact_rationals(dt, name):
    x = _samples(dt, np.random.default_rng(31))
    got = _map(dt, name, x)
    want = np.array([_exact(name, v) for v in x], dtype=dt)
    # values and the sign of zero (a negative input that rounds to 0 gives -0)
    assert np.array_equal(got.view(np.uint32 if dt is np.float32 else np.uint64),
                          want.view(np.uint32 if dt is np.float32 else np.uint64))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_sign_against_comparisons(dt):
    x = _samples(dt, np.random.default_rng(32))
    x = np.concatenate([x, np.array([np.inf, -np.inf, np.finfo(dt).tiny * 1e-3], dtype=dt)])
    got = _map(dt, "sign", x)
    assert np.array_equal(got, np.sign(x))
