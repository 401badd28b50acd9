#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ (run once; outputs committed).

* math_f32.npz / math_f64.npz — inputs and correctly rounded outputs of exp,
  log, erfc, normcdf, cos, sin, exp2, exp10, log2, log10, expm1, log1p, tanh,
  sinh, cosh, tan, asin, acos, erf, cbrt, asinh, acosh, atanh, cot, sec, csc,
  sinpi, cospi, tanpi, cotpi, sind, cosd, tand and cotd, computed with
  mpmath at 60 significant digits.  Inputs: edge values, the NT2 reduction-range boundaries, and seeded
  random points spread over each function's domain (including huge
  arguments for the Payne-Hanek path of cos/sin).
* bs_golden.npz — Black-Scholes call/put in mpmath for 512 options drawn
  from the benchmark distribution (f32 inputs promoted exactly).
* bs64_golden.npz — Black-Scholes call/put of 480 double-precision options
  in mpmath, each price stored as the double-double hi + lo: the benchmark
  distribution, two other rate/volatility pairs, short maturities (deep
  out-of-the-money tails down to ~1e-250) and near-the-money options.
* kats.json — the known-answer tests of NT2's own unit tests for this path,
  transcribed as data (input, expected, tolerance in NT2 ulpdist units, the
  file:line of the reference test).

The reference's tests hold no vector files for this path; its unit tests hold
the KATs below, and the rest is fixed by mathematics (mpmath).
"""
import json
import os

import mpmath
import numpy as np

mpmath.mp.dps = 60
HERE = os.path.dirname(os.path.abspath(__file__))


def round_to(v, dtype):
    """Correctly round the mpmath value v to float32/float64 (ties to even)."""
    if mpmath.isnan(v):
        return dtype(np.nan)
    if mpmath.isinf(v):
        return dtype(np.inf if v > 0 else -np.inf)
    d = float(v)  # correctly rounded double
    if dtype is np.float64:
        return np.float64(d)
    f = np.float32(d)
    if not np.isfinite(f):
        return f
    # the two float32 neighbours bracketing v
    if mpmath.mpf(float(f)) == v:
        return f
    if mpmath.mpf(float(f)) < v:
        lo, hi = f, np.nextafter(f, np.float32(np.inf))
    else:
        lo, hi = np.nextafter(f, np.float32(-np.inf)), f
    dl = v - mpmath.mpf(float(lo))
    dh = mpmath.mpf(float(hi)) - v
    if dl < dh:
        return np.float32(lo)
    if dh < dl:
        return np.float32(hi)
    return np.float32(lo) if (lo.view(np.uint32) & 1) == 0 else np.float32(hi)


FUNCS = {
    "exp": lambda x: mpmath.exp(x),
    "log": lambda x: mpmath.log(x) if x > 0 else (mpmath.mpf("-inf") if x == 0 else mpmath.nan),
    "erfc": lambda x: mpmath.erfc(x),
    "normcdf": lambda x: mpmath.ncdf(x),
    "cos": lambda x: mpmath.cos(x),
    "sin": lambda x: mpmath.sin(x),
    # appended after the first six so their seeded inputs stay as committed
    "exp2": lambda x: mpmath.power(2, x),
    "exp10": lambda x: mpmath.power(10, x),
    "log2": lambda x: mpmath.log(x, 2) if x > 0 else (mpmath.mpf("-inf") if x == 0 else mpmath.nan),
    "log10": lambda x: mpmath.log10(x) if x > 0 else (mpmath.mpf("-inf") if x == 0 else mpmath.nan),
    "expm1": lambda x: mpmath.expm1(x),
    "log1p": lambda x: mpmath.log1p(x) if x > -1 else (mpmath.mpf("-inf") if x == -1 else mpmath.nan),
    "tanh": lambda x: mpmath.tanh(x),
    "sinh": lambda x: mpmath.sinh(x),
    "cosh": lambda x: mpmath.cosh(x),
    "tan": lambda x: mpmath.tan(x),
    "asin": lambda x: mpmath.asin(x) if abs(x) <= 1 else mpmath.nan,
    "acos": lambda x: mpmath.acos(x) if abs(x) <= 1 else mpmath.nan,
    "erf": lambda x: mpmath.erf(x),
    "cbrt": lambda x: mpmath.cbrt(x) if x >= 0 else -mpmath.cbrt(-x),
    "asinh": lambda x: mpmath.asinh(x),
    "acosh": lambda x: mpmath.acosh(x) if x >= 1 else mpmath.nan,
    "atanh": lambda x: (mpmath.atanh(x) if abs(x) < 1 else
                        (mpmath.inf * mpmath.sign(x) if abs(x) == 1 else mpmath.nan)),
    "cot": lambda x: _ratio(mpmath.cos(x), mpmath.sin(x), x),
    "sec": lambda x: 1 / mpmath.cos(x),
    "csc": lambda x: _ratio(mpmath.mpf(1), mpmath.sin(x), x),
    "sinpi": lambda x: mpmath.sinpi(x),
    "cospi": lambda x: mpmath.cospi(x),
    "tanpi": lambda x: _ratio(mpmath.sinpi(x), mpmath.cospi(x), x, zero_is_nan=True),
    "cotpi": lambda x: _ratio(mpmath.cospi(x), mpmath.sinpi(x), x),
    "sind": lambda x: mpmath.sinpi(x / 180),
    "cosd": lambda x: mpmath.cospi(x / 180),
    "tand": lambda x: _ratio(mpmath.sinpi(x / 180), mpmath.cospi(x / 180), x, zero_is_nan=True),
    "cotd": lambda x: _ratio(mpmath.cospi(x / 180), mpmath.sinpi(x / 180), x),
    # the compositions, on domains where NT2's composition is well conditioned
    # (the rounding of its inner argument, 1/x or x/180, moves the result by
    # at most a few ulp there); the rest is checked against the composition
    # with the rounded argument in tests/test_oracle_math.py
    "coth": lambda x: mpmath.coth(x),
    "sech": lambda x: mpmath.sech(x),
    "csch": lambda x: mpmath.csch(x),
    "acoth": lambda x: mpmath.acoth(x),
    "asech": lambda x: mpmath.asech(x),
    "acsch": lambda x: mpmath.acsch(x),
    "acot": lambda x: mpmath.acot(x) if x != 0 else mpmath.pi / 2,
    "asec": lambda x: mpmath.asec(x),
    "acsc": lambda x: mpmath.acsc(x),
    "acotd": lambda x: mpmath.degrees(mpmath.acot(x)) if x != 0 else mpmath.mpf(90),
    "asecd": lambda x: mpmath.degrees(mpmath.asec(x)),
    "acscd": lambda x: mpmath.degrees(mpmath.acsc(x)),
    "acosd": lambda x: mpmath.degrees(mpmath.acos(x)),
    "acospi": lambda x: mpmath.acos(x) / mpmath.pi,
    "asind": lambda x: mpmath.degrees(mpmath.asin(x)),
    "asinpi": lambda x: mpmath.asin(x) / mpmath.pi,
    "atand": lambda x: mpmath.degrees(mpmath.atan(x)),
    "atanpi": lambda x: mpmath.atan(x) / mpmath.pi,
    "indeg": lambda x: mpmath.degrees(x),
    "inrad": lambda x: mpmath.radians(x),
    "secd": lambda x: _ratio(mpmath.mpf(1), mpmath.cospi(x / 180), x, zero_is_nan=True),
    "cscd": lambda x: _ratio(mpmath.mpf(1), mpmath.sinpi(x / 180), x),
    "secpi": lambda x: _ratio(mpmath.mpf(1), mpmath.cospi(x), x, zero_is_nan=True),
    "cscpi": lambda x: _ratio(mpmath.mpf(1), mpmath.sinpi(x), x),
    "sinc": lambda x: mpmath.sin(x) / x if x != 0 else mpmath.mpf(1),
    "sincpi": lambda x: mpmath.sinpi(x) / (mpmath.pi * x) if x != 0 else mpmath.mpf(1),
    "sqrt1pm1": lambda x: mpmath.sqrt(1 + x) - 1 if x >= -1 else mpmath.nan,
    "exprecneg": lambda x: mpmath.exp(-1 / x),
    "exprecnegc": lambda x: -mpmath.expm1(-1 / x),
    "reallog": lambda x: mpmath.log(x),
    "realsqrt": lambda x: mpmath.sqrt(x),
    "expx2": lambda x: mpmath.exp(x * x),
    "erfinv": lambda x: mpmath.erfinv(x) if abs(x) < 1 else (mpmath.inf * mpmath.sign(x) if abs(x) == 1 else mpmath.nan),
    "erfcinv": lambda x: mpmath.erfinv(1 - x) if 0 < x < 2 else (mpmath.inf * mpmath.sign(1 - x) if x in (0, 2) else mpmath.nan),
    "sinhc": lambda x: mpmath.sinh(x) / x if x != 0 else mpmath.mpf(1),
}


def _ratio(num, den, x, zero_is_nan=False):
    """num / den; at an exact zero of den: +-Inf at x = +-0, NaN elsewhere
    (a pole: NT2 returns NaN there, and the sign of the limit is undefined)."""
    if den == 0:
        return mpmath.nan if (zero_is_nan or x != 0) else mpmath.inf * (1 if mpmath.sign(x) >= 0 else -1)
    return num / den


def inputs(name, dtype, rng):
    f = dtype
    big = 3.0e38 if dtype is np.float32 else 1.0e300
    pts = []
    if name == "exp":
        lim = 88.0 if dtype is np.float32 else 708.0
        pts += list(rng.uniform(-lim, lim, 1200)) + list(rng.uniform(-1, 1, 300))
        pts += [0.0, 1.0, -1.0, 0.5, -0.5, lim, -lim, 1e-8, -1e-8]
    elif name in ("exp2", "exp10"):
        # the normal range of the result (NT2 flushes what lies below 2^-126 / 2^-1022)
        lim = (126.0 if dtype is np.float32 else 1021.0) / (1.0 if name == "exp2" else np.log2(10.0))
        pts += list(rng.uniform(-lim, lim, 1200)) + list(rng.uniform(-1, 1, 300))
        pts += [0.0, 1.0, -1.0, 0.5, -0.5, 2.0, 3.0, 10.0, -10.0, 1e-8, -1e-8]
    elif name == "expm1":
        lim = 88.0 if dtype is np.float32 else 709.0
        pts += list(rng.uniform(-20, lim, 800)) + list(rng.uniform(-1, 1, 400))
        pts += list(np.exp2(rng.uniform(-60, -1, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += [0.0, 1.0, -1.0, 0.5, -0.5, 0.34657359, -0.34657359, 1e-8, -1e-8, 10.0]
    elif name == "log1p":
        pts += list(rng.uniform(-0.99, 1, 600)) + list(np.exp2(rng.uniform(-60, 100, 600)))
        pts += list(-np.exp2(rng.uniform(-60, -1, 200)))
        pts += [0.0, 1.0, -0.5, 1e-8, -1e-8, np.e - 1, 1e10]
    elif name == "tanh":
        pts += list(rng.uniform(-0.625, 0.625, 500)) + list(rng.uniform(-10, 10, 600))
        pts += list(np.exp2(rng.uniform(-60, -2, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += list(rng.uniform(10, 25, 100))
        pts += [0.0, 0.625, -0.625, 0.6249, 1.0, -1.0, 5.0, 20.0]
    elif name in ("sinh", "cosh"):
        lim = 89.0 if dtype is np.float32 else 710.0
        pts += list(rng.uniform(-1, 1, 400)) + list(rng.uniform(-lim, lim, 800))
        pts += list(np.exp2(rng.uniform(-60, -1, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += [0.0, 1.0, -1.0, 0.999, 1.001, 20.0, -20.0, lim - 1.0, -(lim - 1.0)]
    elif name in ("log", "log2", "log10"):
        e = rng.uniform(-120 if dtype is np.float32 else -1000, 120 if dtype is np.float32 else 1000, 1200)
        pts += list(np.exp2(e)) + list(rng.uniform(0.5, 2.0, 300))
        pts += [1.0, 2.0, 0.5, np.e, 10.0, 0.7071, 1.4142, 1e-30]
        if dtype is np.float32:
            pts += [1e-40, 3e-45, 3.0e38]
    elif name in ("erfc", "normcdf"):
        pts += list(rng.uniform(-6, 6, 800)) + list(rng.uniform(-1, 1, 300))
        pts += list(rng.uniform(6, 12 if dtype is np.float32 else 26, 200))
        pts += [0.0, 0.5, 1.0, 2.0, 2.0 / 3.0, 0.65, 2.2, 6.0, -0.5, -1.0, -2.0]
        if name == "normcdf":
            pts += list(rng.uniform(-14, -6, 150))
    elif name == "tan":
        pts += list(rng.uniform(-np.pi / 4, np.pi / 4, 300)) + list(rng.uniform(-np.pi / 2, np.pi / 2, 300))
        pts += list(rng.uniform(-20 * np.pi, 20 * np.pi, 300)) + list(rng.uniform(-1e4, 1e4, 300))
        pts += list(np.exp2(rng.uniform(20, 100, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += [0.0, 1.0, -1.0, 0.5, np.pi / 4, 1.5, -1.5, 100.0, 1e5]
    elif name in ("asin", "acos"):
        pts += list(rng.uniform(-1, 1, 800)) + list(rng.uniform(-0.7, 0.7, 300))
        pts += list((1 - np.exp2(rng.uniform(-50, -1, 200))) * rng.choice([-1.0, 1.0], 200))
        pts += list(np.exp2(rng.uniform(-60, -2, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += [0.0, 0.5, -0.5, 0.625, -0.625, 1.0, -1.0, 0.9999, 1e-4, 1.5e-8, 0.25, 0.75]
    elif name == "erf":
        pts += list(rng.uniform(-6.5, 6.5, 800)) + list(rng.uniform(-1, 1, 300))
        pts += list(np.exp2(rng.uniform(-60, -1, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += [0.0, 0.5, 1.0, 2.0, 2.0 / 3.0, 0.65, 2.2, 6.0, -0.5, -1.0, -2.0, 1e-11, 3.0, 5.9]
    elif name == "cbrt":
        lo = -126 if dtype is np.float32 else -1022
        hi = 127 if dtype is np.float32 else 1023
        pts += list(np.exp2(rng.uniform(lo, hi, 900)) * rng.choice([-1.0, 1.0], 900))
        pts += list(rng.uniform(-10, 10, 300))
        pts += [1.0, -1.0, 8.0, -8.0, 27.0, 0.125, 2.0, 4.0, 3.0, 1e-30, 1e30]
    elif name == "asinh":
        lim = 127 if dtype is np.float32 else 1000
        pts += list(rng.uniform(-3, 3, 500)) + list(rng.uniform(-0.5, 0.5, 200))
        pts += list(np.exp2(rng.uniform(-60, lim, 500)) * rng.choice([-1.0, 1.0], 500))
        pts += list(rng.uniform(2000, 4000, 100))
        pts += [0.0, 0.5, -0.5, 0.4999, 1.0, -1.0, 2896.0, 2897.0, 67108864.0, 67108866.0, 1e-9, 1e-20]
    elif name == "acosh":
        lim = 127 if dtype is np.float32 else 1000
        pts += list(1.0 + rng.uniform(0, 3, 500)) + list(1.0 + np.exp2(rng.uniform(-50, -2, 200)))
        pts += list(np.exp2(rng.uniform(1, lim, 500)))
        pts += [1.0, 1.5, 2.0, 10.0, 4194305.0, 4194306.0, 2.0 ** 51 + 2.0 ** 40, 2.0 ** 53]
    elif name == "atanh":
        pts += list(rng.uniform(-1, 1, 800)) + list(rng.uniform(-0.5, 0.5, 300))
        pts += list((1 - np.exp2(rng.uniform(-50, -1, 200))) * rng.choice([-1.0, 1.0], 200))
        pts += list(np.exp2(rng.uniform(-60, -2, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += [0.0, 0.5, -0.5, 0.4999, 0.9, -0.9, 0.25, 1e-9]
    elif name in ("cot", "sec", "csc"):
        pts += list(rng.uniform(-np.pi / 4, np.pi / 4, 300)) + list(rng.uniform(-np.pi / 2, np.pi / 2, 300))
        pts += list(rng.uniform(-20 * np.pi, 20 * np.pi, 300)) + list(rng.uniform(-1e4, 1e4, 300))
        pts += list(np.exp2(rng.uniform(20, 100, 200)) * rng.choice([-1.0, 1.0], 200))
        pts += [1.0, -1.0, 0.5, np.pi / 4, 1.5, -1.5, 100.0, 1e5, 3.0]
    elif name in ("sinpi", "cospi", "tanpi", "cotpi"):
        pts += list(rng.uniform(-4, 4, 800)) + list(rng.uniform(-0.25, 0.25, 300))
        pts += list(np.exp2(rng.uniform(-30, 30, 300)) * rng.choice([-1.0, 1.0], 300))
        pts += list(rng.integers(-50, 50, 60) * 0.25)
        pts += [0.25, -0.25, 0.5, 1.0, 1.5, 2.0, 0.125, 1e-20, 12345.25, 1e7 + 0.5]
    elif name in ("sind", "cosd", "tand", "cotd"):
        pts += list(rng.uniform(-720, 720, 800)) + list(rng.uniform(-45, 45, 300))
        pts += list(np.exp2(rng.uniform(-20, 40, 300)) * rng.choice([-1.0, 1.0], 300))
        pts += list(rng.integers(-64, 64, 60) * 22.5)
        pts += [45.0, -45.0, 90.0, 180.0, 270.0, 30.0, 60.0, 1e-20, 123456.0, 1e7 + 45]
    elif name in COMPOSED_DOMAINS:
        lo, hi, extra = COMPOSED_DOMAINS[name]
        sgn = (lambda k: np.ones(k)) if lo >= 0 else (lambda k: rng.choice([-1.0, 1.0], k))
        a = abs(lo) if lo < 0 else lo
        pts += list(rng.uniform(a, hi, 700) * sgn(700))
        if hi / max(a, 1e-300) > 100:  # log-uniform over wide ranges
            pts += list(np.exp2(rng.uniform(np.log2(max(a, 1e-30)), np.log2(hi), 500)) * sgn(500))
        pts += extra
    else:  # cos, sin
        pts += list(rng.uniform(-np.pi / 4, np.pi / 4, 150))
        pts += list(rng.uniform(-np.pi / 2, np.pi / 2, 150))
        pts += list(rng.uniform(-20 * np.pi, 20 * np.pi, 200))
        pts += list(rng.uniform(-201, 201, 200))
        pts += list(rng.uniform(-1e4, 1e4, 300))
        pts += list(rng.uniform(-823549, 823549, 200))
        pts += list(np.exp2(rng.uniform(20, 127 if dtype is np.float32 else 1020, 300)) *
                    rng.choice([-1.0, 1.0], 300))
        pts += [0.0, np.pi, -np.pi, np.pi / 2, np.pi / 4, 1.0, 100.0, 1e5, 1e10, big]
    x = np.array(pts, dtype=np.float64).astype(dtype)
    return x


# (low |x| or signed low, high |x|, extra points); a negative low means both signs
COMPOSED_DOMAINS = {
    "coth": (-1e-6, 25.0, [0.625, 0.6249, -0.625, 1.0, 5.0, 20.0]),
    "sech": (-1e-6, 80.0, [0.5, 1.0, 0.625, 10.0]),
    "csch": (-1e-6, 80.0, [0.5, 1.0, 0.625, 10.0]),
    "acoth": (-1.25, 1e6, [1.5, 2.0, -2.0, 10.0, 1.6]),
    "asech": (0.001, 0.8, [0.5, 0.25, 0.001]),
    "acsch": (-1e-20, 1e20, [0.5, 1.0, 2.0, -1.0]),
    "acot": (-1e-20, 1e20, [0.5, 1.0, -1.0, 2.0, 0.41421357, 2.4142137]),
    "asec": (-1.25, 1e20, [2.0, -2.0, 1.5, 10.0]),
    "acsc": (-1.25, 1e20, [2.0, -2.0, 1.5, 10.0]),
    "acotd": (-1e-20, 1.5, [0.5, 1.0, -1.0, 1e-3]),
    "asecd": (-1.25, 1e20, [2.0, -2.0, 1.5, 10.0]),
    "acscd": (-1.25, 1e20, [2.0, -2.0, 1.5, 10.0]),
    "acosd": (-1e-20, 1.0, [0.5, -0.5, 1.0, -1.0, 0.0]),
    "acospi": (-1e-20, 1.0, [0.5, -0.5, 1.0, -1.0, 0.0]),
    "asind": (-1e-20, 1.0, [0.5, -0.5, 1.0, -1.0]),
    "asinpi": (-1e-20, 1.0, [0.5, -0.5, 1.0, -1.0]),
    "atand": (-1e-20, 1e20, [0.5, 1.0, -1.0, 2.0]),
    "atanpi": (-1e-20, 1e20, [0.5, 1.0, -1.0, 2.0]),
    "indeg": (-1e-30, 1e30, [np.pi, np.pi / 2, np.pi / 4, 1.0]),
    "inrad": (-1e-30, 1e30, [180.0, 90.0, 45.0, 1.0]),
    "secd": (-0.001, 720.0, [0.0, 45.0, 60.0, 90.0, 180.0, 270.0, -45.0]),
    "cscd": (-0.001, 120.0, [45.0, 90.0, 30.0, -45.0, 1e-20]),
    "secpi": (-1e-6, 4.0, [0.0, 0.25, 1.0, 0.5, 1.5, 1.0 / 3.0]),
    "cscpi": (-1e-6, 4.0, [0.25, 0.5, 1.0, -0.25, 1.0 / 6.0]),
    "sinc": (-1e-20, 60.0, [np.pi / 2, np.pi / 4, 1.0, 1e-4, 3.0]),
    "sincpi": (-1e-20, 20.0, [0.5, 0.25, 1.0, 1e-4, 1.5]),
    "sqrt1pm1": (-1e-20, 1.0, [0.5, -0.5, 0.49999, 2.0, 10.0, 1e10, 1e30, -1.0]),
    "exprecneg": (-0.5, 1e20, [1.0, -1.0, 2.0, 0.5]),
    "exprecnegc": (-0.5, 1e20, [1.0, -1.0, 2.0, 0.5]),
    "reallog": (1e-30, 1e30, [1.0, 2.0, 0.5, 10.0]),
    "realsqrt": (1e-30, 1e30, [1.0, 2.0, 4.0, 0.25]),
    "expx2": (-1e-6, 9.0, [1.0, 0.5, 3.0, 5.0, 8.9, 0.01]),
    "erfinv": (-1e-20, 1.0, [0.5, -0.5, 0.9966, 0.99, 0.999, 0.9999999, -0.99999, 1e-5]),
    "erfcinv": (1e-30, 2.0, [1.0, 0.5, 1.5, 0.0034, 1.9966, 1e-10, 1e-20, 1.9999999]),
    "sinhc": (-1e-6, 85.0, [1.0, 0.999, 0.5, 2.0, 10.0, 20.0]),
}


def build_math(dtype, tag):
    rng = np.random.default_rng(20240601 if dtype is np.float32 else 20240602)
    out = {}
    for name, fn in FUNCS.items():
        x = inputs(name, dtype, rng)
        y = np.array([round_to(fn(mpmath.mpf(float(v))), dtype) for v in x], dtype=dtype)
        out[name + "_x"] = x
        out[name + "_y"] = y
    np.savez(os.path.join(HERE, "math_%s.npz" % tag), **out)


def build_bs():
    rng = np.random.default_rng(2)
    n = 512
    S = rng.uniform(5, 30, n).astype(np.float32)
    X = rng.uniform(1, 100, n).astype(np.float32)
    T = rng.uniform(0.25, 10, n).astype(np.float32)
    r, v = mpmath.mpf("0.02"), mpmath.mpf("0.30")
    call = np.empty(n)
    put = np.empty(n)
    for i in range(n):
        s, x, t = (mpmath.mpf(float(a)) for a in (S[i], X[i], T[i]))
        vst = v * mpmath.sqrt(t)
        d1 = (mpmath.log(s / x) + (r + v * v / 2) * t) / vst
        d2 = d1 - vst
        e = x * mpmath.exp(-r * t)
        call[i] = float(s * mpmath.ncdf(d1) - e * mpmath.ncdf(d2))
        put[i] = float(e * mpmath.ncdf(-d2) - s * mpmath.ncdf(-d1))
    np.savez(os.path.join(HERE, "bs_golden.npz"), S=S, X=X, T=T, r=np.float64(0.02),
             v=np.float64(0.30), call=call, put=put)


def _dd(v):
    hi = float(v)
    return hi, float(v - mpmath.mpf(hi))


def build_bs64():
    rng = np.random.default_rng(64)
    groups = [  # (count, r, v, S range, X range, T range)
        (160, 0.02, 0.30, (5, 30), (1, 100), (0.25, 10)),
        (80, 0.09, 0.30, (5, 30), (1, 100), (0.25, 10)),
        (80, -0.012, 0.45, (5, 30), (1, 100), (0.25, 30)),
        (80, 0.02, 0.30, (5, 30), (1, 100), (0.002, 0.25)),
        (80, 0.02, 0.30, (10, 20), (9.5, 21), (0.25, 10)),
    ]
    cols = {k: [] for k in ("S", "X", "T", "r", "v", "call_hi", "call_lo", "put_hi", "put_lo")}
    for cnt, r, v, sr, xr, tr in groups:
        S = rng.uniform(*sr, cnt)
        X = rng.uniform(*xr, cnt)
        T = rng.uniform(*tr, cnt)
        rm, vm = mpmath.mpf(r), mpmath.mpf(v)
        for i in range(cnt):
            s, x, t = (mpmath.mpf(float(a)) for a in (S[i], X[i], T[i]))
            vst = vm * mpmath.sqrt(t)
            d1 = (mpmath.log(s / x) + (rm + vm * vm / 2) * t) / vst
            d2 = d1 - vst
            e = x * mpmath.exp(-rm * t)
            c = s * mpmath.ncdf(d1) - e * mpmath.ncdf(d2)
            pt = e * mpmath.ncdf(-d2) - s * mpmath.ncdf(-d1)
            for k, val in (("S", S[i]), ("X", X[i]), ("T", T[i]), ("r", r), ("v", v)):
                cols[k].append(float(val))
            for k, val in (("call", c), ("put", pt)):
                h, lo = _dd(val)
                cols[k + "_hi"].append(h)
                cols[k + "_lo"].append(lo)
    np.savez(os.path.join(HERE, "bs64_golden.npz"), **{k: np.array(v) for k, v in cols.items()})


# Known answers from NT2's unit tests (value, expected, tolerance in ulpdist)
KATS = [
    # modules/core/exponential/unit/simd/exp.cpp:42-49
    ("exp", "inf", "inf", 0.75, "modules/core/exponential/unit/simd/exp.cpp:42"),
    ("exp", "-inf", "0", 0.75, "modules/core/exponential/unit/simd/exp.cpp:43"),
    ("exp", "nan", "nan", 0.75, "modules/core/exponential/unit/simd/exp.cpp:44"),
    ("exp", "-1", "1/e", 0.75, "modules/core/exponential/unit/simd/exp.cpp:46"),
    ("exp", "1", "e", 0.75, "modules/core/exponential/unit/simd/exp.cpp:47"),
    ("exp", "0", "1", 0.75, "modules/core/exponential/unit/simd/exp.cpp:48"),
    # modules/core/exponential/unit/simd/log.cpp:42-50
    ("log", "inf", "inf", 0, "modules/core/exponential/unit/simd/log.cpp:42"),
    ("log", "-inf", "nan", 0, "modules/core/exponential/unit/simd/log.cpp:43"),
    ("log", "nan", "nan", 0, "modules/core/exponential/unit/simd/log.cpp:44"),
    ("log", "-1", "nan", 0, "modules/core/exponential/unit/simd/log.cpp:45"),
    ("log", "0", "-inf", 0, "modules/core/exponential/unit/simd/log.cpp:46"),
    ("log", "1", "0", 0, "modules/core/exponential/unit/simd/log.cpp:48"),
    ("log", "e", "1", 1, "modules/core/exponential/unit/simd/log.cpp:49"),
    # modules/core/euler/unit/simd/erfc.cpp:45-54
    ("erfc", "inf", "0", 200, "modules/core/euler/unit/simd/erfc.cpp:45"),
    ("erfc", "nan", "nan", 200, "modules/core/euler/unit/simd/erfc.cpp:46"),
    ("erfc", "-0", "1", 200, "modules/core/euler/unit/simd/erfc.cpp:48"),
    ("erfc", "1", "0.157299207050285", 200, "modules/core/euler/unit/simd/erfc.cpp:49"),
    ("erfc", "0.5", "0.479500122186953", 200, "modules/core/euler/unit/simd/erfc.cpp:50"),
    ("erfc", "2", "0.004677734981047", 200, "modules/core/euler/unit/simd/erfc.cpp:51"),
    ("erfc", "10", "0", 200, "modules/core/euler/unit/simd/erfc.cpp:52"),
    ("erfc", "0", "1", 200, "modules/core/euler/unit/simd/erfc.cpp:53"),
    # modules/core/statistics/unit/scalar/normcdf.cpp:31-36
    ("normcdf", "nan", "nan", 0.5, "modules/core/statistics/unit/scalar/normcdf.cpp:31"),
    ("normcdf", "1", "0.841344746068542948585232545632", 1.5, "modules/core/statistics/unit/scalar/normcdf.cpp:32"),
    ("normcdf", "-1", "0.158655253931457051414767454368", 6, "modules/core/statistics/unit/scalar/normcdf.cpp:33"),
    ("normcdf", "0", "0.5", 0.5, "modules/core/statistics/unit/scalar/normcdf.cpp:34"),
    ("normcdf", "inf", "1", 0.5, "modules/core/statistics/unit/scalar/normcdf.cpp:35"),
    ("normcdf", "-inf", "0", 0.5, "modules/core/statistics/unit/scalar/normcdf.cpp:36"),
    # normcdf(-1, 1) = normcdf(-2): normcdf.cpp:49 (the two-argument form shifts by mu)
    ("normcdf", "-2", "0.0227501319481792072002826371665", 25, "modules/core/statistics/unit/scalar/normcdf.cpp:49"),
    # modules/core/trigonometric/unit/simd/cos.cpp:50-60
    ("cos", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:50"),
    ("cos", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:51"),
    ("cos", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:52"),
    ("cos", "-pi", "-1", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:54"),
    ("cos", "-pi/2", "0", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:55"),
    ("cos", "-pi/4", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:56"),
    ("cos", "pi", "-1", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:57"),
    ("cos", "pi/2", "0", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:58"),
    ("cos", "pi/4", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:59"),
    ("cos", "0", "1", 0.5, "modules/core/trigonometric/unit/simd/cos.cpp:60"),
    # modules/core/exponential/unit/simd/exp2.cpp:46-52
    ("exp2", "inf", "inf", 0, "modules/core/exponential/unit/simd/exp2.cpp:46"),
    ("exp2", "-inf", "0", 0, "modules/core/exponential/unit/simd/exp2.cpp:47"),
    ("exp2", "nan", "nan", 0, "modules/core/exponential/unit/simd/exp2.cpp:48"),
    ("exp2", "-1", "0.5", 0, "modules/core/exponential/unit/simd/exp2.cpp:50"),
    ("exp2", "1", "2", 0, "modules/core/exponential/unit/simd/exp2.cpp:51"),
    ("exp2", "0", "1", 0, "modules/core/exponential/unit/simd/exp2.cpp:52"),
    # modules/core/exponential/unit/simd/exp10.cpp:46-52
    ("exp10", "inf", "inf", 0, "modules/core/exponential/unit/simd/exp10.cpp:46"),
    ("exp10", "-inf", "0", 0, "modules/core/exponential/unit/simd/exp10.cpp:47"),
    ("exp10", "nan", "nan", 0, "modules/core/exponential/unit/simd/exp10.cpp:48"),
    ("exp10", "-1", "0.1", 0, "modules/core/exponential/unit/simd/exp10.cpp:50"),
    ("exp10", "1", "10", 0, "modules/core/exponential/unit/simd/exp10.cpp:51"),
    ("exp10", "0", "1", 0, "modules/core/exponential/unit/simd/exp10.cpp:52"),
    # modules/core/exponential/unit/simd/log2.cpp:43-49
    ("log2", "inf", "inf", 0, "modules/core/exponential/unit/simd/log2.cpp:43"),
    ("log2", "-inf", "nan", 0, "modules/core/exponential/unit/simd/log2.cpp:44"),
    ("log2", "nan", "nan", 0, "modules/core/exponential/unit/simd/log2.cpp:45"),
    ("log2", "-1", "nan", 0, "modules/core/exponential/unit/simd/log2.cpp:46"),
    ("log2", "0", "-inf", 0, "modules/core/exponential/unit/simd/log2.cpp:47"),
    ("log2", "1", "0", 0, "modules/core/exponential/unit/simd/log2.cpp:49"),
    # modules/core/exponential/unit/simd/log10.cpp:46-52
    ("log10", "inf", "inf", 0, "modules/core/exponential/unit/simd/log10.cpp:46"),
    ("log10", "-inf", "nan", 0, "modules/core/exponential/unit/simd/log10.cpp:47"),
    ("log10", "nan", "nan", 0, "modules/core/exponential/unit/simd/log10.cpp:48"),
    ("log10", "-1", "nan", 0, "modules/core/exponential/unit/simd/log10.cpp:49"),
    ("log10", "0", "-inf", 0, "modules/core/exponential/unit/simd/log10.cpp:50"),
    ("log10", "1", "0", 0, "modules/core/exponential/unit/simd/log10.cpp:52"),
    # modules/core/exponential/unit/simd/expm1.cpp:44-50
    ("expm1", "inf", "inf", 1.0, "modules/core/exponential/unit/simd/expm1.cpp:44"),
    ("expm1", "-inf", "-1", 1.0, "modules/core/exponential/unit/simd/expm1.cpp:45"),
    ("expm1", "nan", "nan", 1.0, "modules/core/exponential/unit/simd/expm1.cpp:46"),
    ("expm1", "-1", "1/e-1", 1.0, "modules/core/exponential/unit/simd/expm1.cpp:48"),
    ("expm1", "1", "e-1", 1.0, "modules/core/exponential/unit/simd/expm1.cpp:49"),
    ("expm1", "0", "0", 1.0, "modules/core/exponential/unit/simd/expm1.cpp:50"),
    # modules/core/exponential/unit/simd/log1p.cpp:46-54
    ("log1p", "inf", "inf", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:46"),
    ("log1p", "-inf", "nan", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:47"),
    ("log1p", "nan", "nan", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:48"),
    ("log1p", "-1", "-inf", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:49"),
    ("log1p", "eps", "eps", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:51"),
    ("log1p", "1", "ln2", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:52"),
    ("log1p", "0", "0", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:53"),
    ("log1p", "tiny", "tiny", 0.5, "modules/core/exponential/unit/simd/log1p.cpp:54"),
    # modules/core/hyperbolic/unit/simd/tanh.cpp:43-49
    ("tanh", "inf", "1", 0.5, "modules/core/hyperbolic/unit/simd/tanh.cpp:43"),
    ("tanh", "-inf", "-1", 0.5, "modules/core/hyperbolic/unit/simd/tanh.cpp:44"),
    ("tanh", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/simd/tanh.cpp:45"),
    ("tanh", "-1", "-0.7615941559557649", 0.5, "modules/core/hyperbolic/unit/simd/tanh.cpp:47"),
    ("tanh", "1", "0.7615941559557649", 0.5, "modules/core/hyperbolic/unit/simd/tanh.cpp:48"),
    ("tanh", "0", "0", 0.5, "modules/core/hyperbolic/unit/simd/tanh.cpp:49"),
    # modules/core/hyperbolic/unit/simd/sinh.cpp:43-47 and cosh.cpp:42-46
    ("sinh", "inf", "inf", 0.5, "modules/core/hyperbolic/unit/simd/sinh.cpp:43"),
    ("sinh", "-inf", "-inf", 0.5, "modules/core/hyperbolic/unit/simd/sinh.cpp:44"),
    ("sinh", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/simd/sinh.cpp:45"),
    ("sinh", "0", "0", 0.5, "modules/core/hyperbolic/unit/simd/sinh.cpp:47"),
    ("cosh", "inf", "inf", 0.5, "modules/core/hyperbolic/unit/simd/cosh.cpp:42"),
    ("cosh", "-inf", "inf", 0.5, "modules/core/hyperbolic/unit/simd/cosh.cpp:43"),
    ("cosh", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/simd/cosh.cpp:44"),
    ("cosh", "0", "1", 0.5, "modules/core/hyperbolic/unit/simd/cosh.cpp:46"),
    # modules/core/trigonometric/unit/simd/tan.cpp:45-53
    ("tan", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/tan.cpp:45"),
    ("tan", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/tan.cpp:46"),
    ("tan", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/tan.cpp:47"),
    ("tan", "-pi", "0", 0.75, "modules/core/trigonometric/unit/simd/tan.cpp:49"),
    ("tan", "-pi/4", "-1", 1.0, "modules/core/trigonometric/unit/simd/tan.cpp:50"),
    ("tan", "pi", "0", 0.75, "modules/core/trigonometric/unit/simd/tan.cpp:51"),
    ("tan", "pi/4", "1", 1.0, "modules/core/trigonometric/unit/simd/tan.cpp:52"),
    ("tan", "0", "0", 0.5, "modules/core/trigonometric/unit/simd/tan.cpp:53"),
    # modules/core/trigonometric/unit/simd/asin.cpp:49-57 and acos.cpp:49-57
    ("asin", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:49"),
    ("asin", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:50"),
    ("asin", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:51"),
    ("asin", "0.5", "pi/6", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:53"),
    ("asin", "-0.5", "-pi/6", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:54"),
    ("asin", "-1", "-pi/2", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:55"),
    ("asin", "1", "pi/2", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:56"),
    ("asin", "0", "0", 0.5, "modules/core/trigonometric/unit/simd/asin.cpp:57"),
    ("acos", "inf", "nan", 0.0, "modules/core/trigonometric/unit/simd/acos.cpp:49"),
    ("acos", "-inf", "nan", 0.0, "modules/core/trigonometric/unit/simd/acos.cpp:50"),
    ("acos", "nan", "nan", 0.0, "modules/core/trigonometric/unit/simd/acos.cpp:51"),
    ("acos", "0.5", "pi/3", 0.5, "modules/core/trigonometric/unit/simd/acos.cpp:53"),
    ("acos", "-0.5", "2pi/3", 0.5, "modules/core/trigonometric/unit/simd/acos.cpp:54"),
    ("acos", "-1", "pi", 0.0, "modules/core/trigonometric/unit/simd/acos.cpp:55"),
    ("acos", "1", "0", 0.0, "modules/core/trigonometric/unit/simd/acos.cpp:56"),
    ("acos", "0", "pi/2", 0.0, "modules/core/trigonometric/unit/simd/acos.cpp:57"),
    # modules/core/euler/unit/simd/erf.cpp:45-53
    ("erf", "inf", "1", 10, "modules/core/euler/unit/simd/erf.cpp:45"),
    ("erf", "nan", "nan", 10, "modules/core/euler/unit/simd/erf.cpp:46"),
    ("erf", "-0", "0", 10, "modules/core/euler/unit/simd/erf.cpp:48"),
    ("erf", "1", "0.842700792949715", 10, "modules/core/euler/unit/simd/erf.cpp:49"),
    ("erf", "0.5", "0.520499877813047", 10, "modules/core/euler/unit/simd/erf.cpp:50"),
    ("erf", "2", "0.995322265018953", 10, "modules/core/euler/unit/simd/erf.cpp:51"),
    ("erf", "10", "1", 10, "modules/core/euler/unit/simd/erf.cpp:52"),
    ("erf", "0", "0", 10, "modules/core/euler/unit/simd/erf.cpp:53"),
    # modules/core/exponential/unit/simd/cbrt.cpp:43-50
    ("cbrt", "inf", "inf", 0.5, "modules/core/exponential/unit/simd/cbrt.cpp:43"),
    ("cbrt", "-inf", "-inf", 0.5, "modules/core/exponential/unit/simd/cbrt.cpp:44"),
    ("cbrt", "nan", "nan", 0.5, "modules/core/exponential/unit/simd/cbrt.cpp:45"),
    ("cbrt", "8", "2", 0.5, "modules/core/exponential/unit/simd/cbrt.cpp:47"),
    ("cbrt", "-1", "-1", 0.5, "modules/core/exponential/unit/simd/cbrt.cpp:48"),
    ("cbrt", "1", "1", 0.5, "modules/core/exponential/unit/simd/cbrt.cpp:49"),
    ("cbrt", "0", "0", 0.5, "modules/core/exponential/unit/simd/cbrt.cpp:50"),
    # modules/core/hyperbolic/unit/simd/asinh.cpp:50-57 (the Valmax and 2/sqrt(eps)
    # rows compare with boost::math; their mpmath values are in the golden vectors)
    ("asinh", "inf", "inf", 0.5, "modules/core/hyperbolic/unit/simd/asinh.cpp:50"),
    ("asinh", "-inf", "-inf", 0.5, "modules/core/hyperbolic/unit/simd/asinh.cpp:51"),
    ("asinh", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/simd/asinh.cpp:52"),
    ("asinh", "0", "0", 0.5, "modules/core/hyperbolic/unit/simd/asinh.cpp:54"),
    ("asinh", "eps", "eps", 0.5, "modules/core/hyperbolic/unit/simd/asinh.cpp:57"),
    # modules/core/hyperbolic/unit/simd/acosh.cpp:41-46
    ("acosh", "inf", "inf", 0.5, "modules/core/hyperbolic/unit/simd/acosh.cpp:41"),
    ("acosh", "-inf", "nan", 0.5, "modules/core/hyperbolic/unit/simd/acosh.cpp:42"),
    ("acosh", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/simd/acosh.cpp:43"),
    ("acosh", "0", "nan", 0.5, "modules/core/hyperbolic/unit/simd/acosh.cpp:44"),
    ("acosh", "1", "0", 0.5, "modules/core/hyperbolic/unit/simd/acosh.cpp:46"),
    # modules/core/hyperbolic/unit/simd/atanh.cpp:43-49
    ("atanh", "inf", "nan", 0.5, "modules/core/hyperbolic/unit/simd/atanh.cpp:43"),
    ("atanh", "-inf", "nan", 0.5, "modules/core/hyperbolic/unit/simd/atanh.cpp:44"),
    ("atanh", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/simd/atanh.cpp:45"),
    ("atanh", "-1", "-inf", 0.5, "modules/core/hyperbolic/unit/simd/atanh.cpp:46"),
    ("atanh", "1", "inf", 0.5, "modules/core/hyperbolic/unit/simd/atanh.cpp:47"),
    ("atanh", "0", "0", 0.5, "modules/core/hyperbolic/unit/simd/atanh.cpp:49"),
    # modules/core/trigonometric/unit/simd/cot.cpp:45-54
    ("cot", "-0", "-inf", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:45"),
    ("cot", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:46"),
    ("cot", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:47"),
    ("cot", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:48"),
    ("cot", "0", "inf", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:49"),
    ("cot", "-pi/2", "0", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:51"),
    ("cot", "-pi/4", "-1", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:52"),
    ("cot", "pi/2", "0", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:53"),
    ("cot", "pi/4", "1", 0.5, "modules/core/trigonometric/unit/simd/cot.cpp:54"),
    # modules/core/trigonometric/unit/simd/sec.cpp:46-54 and csc.cpp:46-55
    ("sec", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:46"),
    ("sec", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:47"),
    ("sec", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:48"),
    ("sec", "-pi", "-1", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:50"),
    ("sec", "-pi/4", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:51"),
    ("sec", "pi", "-1", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:52"),
    ("sec", "pi/4", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:53"),
    ("sec", "0", "1", 0.5, "modules/core/trigonometric/unit/simd/sec.cpp:54"),
    ("csc", "-0", "-inf", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:46"),
    ("csc", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:47"),
    ("csc", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:48"),
    ("csc", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:49"),
    ("csc", "0", "inf", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:50"),
    ("csc", "-pi/2", "-1", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:52"),
    ("csc", "-pi/4", "-sqrt(2)", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:53"),
    ("csc", "pi/2", "1", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:54"),
    ("csc", "pi/4", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/simd/csc.cpp:55"),
    # modules/core/trigonometric/unit/simd/sinpi.cpp:48-57, cospi.cpp:49-58
    ("sinpi", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:48"),
    ("sinpi", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:49"),
    ("sinpi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:50"),
    ("sinpi", "-0.25", "-sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:52"),
    ("sinpi", "0.5", "1", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:53"),
    ("sinpi", "-0.5", "-1", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:54"),
    ("sinpi", "1", "0", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:55"),
    ("sinpi", "0.25", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:56"),
    ("sinpi", "0", "0", 0.5, "modules/core/trigonometric/unit/simd/sinpi.cpp:57"),
    ("cospi", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:49"),
    ("cospi", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:50"),
    ("cospi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:51"),
    ("cospi", "-0.25", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:53"),
    ("cospi", "0.5", "0", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:54"),
    ("cospi", "-0.5", "0", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:55"),
    ("cospi", "1", "-1", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:56"),
    ("cospi", "0.25", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:57"),
    ("cospi", "0", "1", 0.5, "modules/core/trigonometric/unit/simd/cospi.cpp:58"),
    # modules/core/trigonometric/unit/simd/tanpi.cpp:47-57, cotpi.cpp:47-57
    ("tanpi", "0.5", "nan", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:47"),
    ("tanpi", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:48"),
    ("tanpi", "-0.5", "nan", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:49"),
    ("tanpi", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:50"),
    ("tanpi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:51"),
    ("tanpi", "-0.25", "-1", 1.0, "modules/core/trigonometric/unit/simd/tanpi.cpp:53"),
    ("tanpi", "-1", "0", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:54"),
    ("tanpi", "1", "0", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:55"),
    ("tanpi", "0.25", "1", 1.0, "modules/core/trigonometric/unit/simd/tanpi.cpp:56"),
    ("tanpi", "0", "0", 0.5, "modules/core/trigonometric/unit/simd/tanpi.cpp:57"),
    ("cotpi", "-0", "-inf", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:47"),
    ("cotpi", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:48"),
    ("cotpi", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:49"),
    ("cotpi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:50"),
    ("cotpi", "1", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:51"),
    ("cotpi", "0", "inf", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:52"),
    ("cotpi", "-0.25", "-1", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:54"),
    ("cotpi", "0.5", "0", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:55"),
    ("cotpi", "-0.5", "0", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:56"),
    ("cotpi", "0.25", "1", 0.5, "modules/core/trigonometric/unit/simd/cotpi.cpp:57"),
    # modules/core/trigonometric/unit/simd/sind.cpp:48-58, cosd.cpp:46-56
    ("sind", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:48"),
    ("sind", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:49"),
    ("sind", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:50"),
    ("sind", "-180", "0", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:52"),
    ("sind", "-45", "-sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:53"),
    ("sind", "-90", "-1", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:54"),
    ("sind", "0", "0", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:55"),
    ("sind", "180", "0", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:56"),
    ("sind", "45", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:57"),
    ("sind", "90", "1", 0.5, "modules/core/trigonometric/unit/simd/sind.cpp:58"),
    ("cosd", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:46"),
    ("cosd", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:47"),
    ("cosd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:48"),
    ("cosd", "-180", "-1", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:50"),
    ("cosd", "-45", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:51"),
    ("cosd", "-90", "0", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:52"),
    ("cosd", "180", "-1", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:53"),
    ("cosd", "45", "sqrt(2)/2", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:54"),
    ("cosd", "90", "0", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:55"),
    ("cosd", "0", "1", 0.5, "modules/core/trigonometric/unit/simd/cosd.cpp:56"),
    # modules/core/trigonometric/unit/simd/tand.cpp:47-57, cotd.cpp:47-58
    ("tand", "-90", "nan", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:47"),
    ("tand", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:48"),
    ("tand", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:49"),
    ("tand", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:50"),
    ("tand", "90", "nan", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:51"),
    ("tand", "-180", "0", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:53"),
    ("tand", "-45", "-1", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:54"),
    ("tand", "0", "0", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:55"),
    ("tand", "180", "0", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:56"),
    ("tand", "45", "1", 0.5, "modules/core/trigonometric/unit/simd/tand.cpp:57"),
    ("cotd", "-0", "-inf", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:47"),
    ("cotd", "-180", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:48"),
    ("cotd", "inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:49"),
    ("cotd", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:50"),
    ("cotd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:51"),
    ("cotd", "0", "inf", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:52"),
    ("cotd", "180", "nan", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:53"),
    ("cotd", "-45", "-1", 4.0, "modules/core/trigonometric/unit/simd/cotd.cpp:55"),
    ("cotd", "-90", "0", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:56"),
    ("cotd", "45", "1", 4.0, "modules/core/trigonometric/unit/simd/cotd.cpp:57"),
    ("cotd", "90", "0", 0.5, "modules/core/trigonometric/unit/simd/cotd.cpp:58"),
    ("sqrt1pm1", "inf", "inf", 0, "modules/core/exponential/unit/scalar/sqrt1pm1.cpp:41"),
    ("sqrt1pm1", "-inf", "nan", 0, "modules/core/exponential/unit/scalar/sqrt1pm1.cpp:42"),
    ("sqrt1pm1", "nan", "nan", 0, "modules/core/exponential/unit/scalar/sqrt1pm1.cpp:43"),
    ("sqrt1pm1", "-1", "-1", 0, "modules/core/exponential/unit/scalar/sqrt1pm1.cpp:45"),
    ("sqrt1pm1", "1", "0.41421356237309503", 2, "modules/core/exponential/unit/scalar/sqrt1pm1.cpp:46"),
    ("sqrt1pm1", "0", "0", 0, "modules/core/exponential/unit/scalar/sqrt1pm1.cpp:47"),
    ("exprecneg", "inf", "1", 0.75, "modules/core/exponential/unit/scalar/exprecneg.cpp:41"),
    ("exprecneg", "-inf", "1", 0.75, "modules/core/exponential/unit/scalar/exprecneg.cpp:42"),
    ("exprecneg", "nan", "nan", 0.75, "modules/core/exponential/unit/scalar/exprecneg.cpp:43"),
    ("exprecneg", "-1", "e", 0.75, "modules/core/exponential/unit/scalar/exprecneg.cpp:45"),
    ("exprecneg", "1", "1/e", 0.75, "modules/core/exponential/unit/scalar/exprecneg.cpp:46"),
    ("exprecneg", "0", "0", 0.75, "modules/core/exponential/unit/scalar/exprecneg.cpp:47"),
    ("exprecneg", "-0", "inf", 0.75, "modules/core/exponential/unit/scalar/exprecneg.cpp:48"),
    ("exprecnegc", "inf", "0", 0.75, "modules/core/exponential/unit/scalar/exprecnegc.cpp:42"),
    ("exprecnegc", "-inf", "0", 0.75, "modules/core/exponential/unit/scalar/exprecnegc.cpp:43"),
    ("exprecnegc", "nan", "nan", 0.75, "modules/core/exponential/unit/scalar/exprecnegc.cpp:44"),
    ("exprecnegc", "-1", "-1.718281828459045", 0.75, "modules/core/exponential/unit/scalar/exprecnegc.cpp:46"),
    ("exprecnegc", "1", "0.6321205588285577", 0.75, "modules/core/exponential/unit/scalar/exprecnegc.cpp:48"),
    ("exprecnegc", "0", "1", 0.75, "modules/core/exponential/unit/scalar/exprecnegc.cpp:50"),
    ("exprecnegc", "-0", "-inf", 0.75, "modules/core/exponential/unit/scalar/exprecnegc.cpp:51"),
    ("reallog", "inf", "inf", 0, "modules/core/exponential/unit/scalar/reallog.cpp:40"),
    ("reallog", "nan", "nan", 0, "modules/core/exponential/unit/scalar/reallog.cpp:41"),
    ("reallog", "0", "-inf", 0, "modules/core/exponential/unit/scalar/reallog.cpp:42"),
    ("reallog", "1", "0", 0, "modules/core/exponential/unit/scalar/reallog.cpp:44"),
    ("realsqrt", "inf", "inf", 0, "modules/core/exponential/unit/scalar/realsqrt.cpp:39"),
    ("realsqrt", "nan", "nan", 0, "modules/core/exponential/unit/scalar/realsqrt.cpp:40"),
    ("realsqrt", "1", "1", 0, "modules/core/exponential/unit/scalar/realsqrt.cpp:42"),
    ("realsqrt", "0", "0", 0, "modules/core/exponential/unit/scalar/realsqrt.cpp:43"),
    ("expx2", "inf", "inf", 1.0, "modules/core/exponential/unit/scalar/expx2.cpp:40"),
    ("expx2", "-inf", "inf", 1.0, "modules/core/exponential/unit/scalar/expx2.cpp:41"),
    ("expx2", "nan", "nan", 1.0, "modules/core/exponential/unit/scalar/expx2.cpp:42"),
    ("expx2", "-1", "e", 1.0, "modules/core/exponential/unit/scalar/expx2.cpp:44"),
    ("expx2", "1", "e", 1.0, "modules/core/exponential/unit/scalar/expx2.cpp:45"),
    ("expx2", "0", "1", 1.0, "modules/core/exponential/unit/scalar/expx2.cpp:46"),
    ("erfinv", "nan", "nan", 2, "modules/core/euler/unit/scalar/erfinv.cpp:40"),
    ("erfinv", "0.52049987781304653768274665389196", "0.5", 2, "modules/core/euler/unit/scalar/erfinv.cpp:42"),
    ("erfinv", "0.84270079294971486934122063508262", "1", 2, "modules/core/euler/unit/scalar/erfinv.cpp:43"),
    ("erfinv", "0.99532226501895273416206925636725", "2", 2, "modules/core/euler/unit/scalar/erfinv.cpp:44"),
    ("erfinv", "0", "0", 2, "modules/core/euler/unit/scalar/erfinv.cpp:45"),
    ("erfinv", "1", "inf", 2, "modules/core/euler/unit/scalar/erfinv.cpp:46"),
    ("erfinv", "-1", "-inf", 2, "modules/core/euler/unit/scalar/erfinv.cpp:47"),
    ("erfcinv", "nan", "nan", 10, "modules/core/euler/unit/scalar/erfcinv.cpp:39"),
    ("erfcinv", "0", "inf", 10, "modules/core/euler/unit/scalar/erfcinv.cpp:40"),
    ("erfcinv", "1", "-0", 10, "modules/core/euler/unit/scalar/erfcinv.cpp:42"),
    ("erfcinv", "0.47950012218695346231725334610804", "0.5", 10, "modules/core/euler/unit/scalar/erfcinv.cpp:43"),
    ("erfcinv", "0.15729920705028513065877936491739", "1", 10, "modules/core/euler/unit/scalar/erfcinv.cpp:44"),
    ("erfcinv", "0.00467773498104726583793074363275", "2", 10, "modules/core/euler/unit/scalar/erfcinv.cpp:45"),
    ("sinhc", "inf", "inf", 0.5, "modules/core/hyperbolic/unit/scalar/sinhc.cpp:37"),
    ("sinhc", "-inf", "inf", 0.5, "modules/core/hyperbolic/unit/scalar/sinhc.cpp:38"),
    ("sinhc", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/sinhc.cpp:39"),
    ("sinhc", "0", "1", 0.5, "modules/core/hyperbolic/unit/scalar/sinhc.cpp:41"),
    # compositions: the scalar unit tests of NT2 (evaluated in packs and as tails)
    ("coth", "inf", "1", 0.5, "modules/core/hyperbolic/unit/scalar/coth.cpp:39"),
    ("coth", "-inf", "-1", 0.5, "modules/core/hyperbolic/unit/scalar/coth.cpp:40"),
    ("coth", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/coth.cpp:41"),
    ("coth", "-0", "-inf", 0.5, "modules/core/hyperbolic/unit/scalar/coth.cpp:42"),
    ("coth", "0", "inf", 0.5, "modules/core/hyperbolic/unit/scalar/coth.cpp:43"),
    ("sech", "inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/sech.cpp:37"),
    ("sech", "-inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/sech.cpp:38"),
    ("sech", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/sech.cpp:39"),
    ("sech", "0", "1", 0.5, "modules/core/hyperbolic/unit/scalar/sech.cpp:41"),
    ("csch", "-0", "-inf", 0.5, "modules/core/hyperbolic/unit/scalar/csch.cpp:36"),
    ("csch", "inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/csch.cpp:37"),
    ("csch", "-inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/csch.cpp:38"),
    ("csch", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/csch.cpp:39"),
    ("csch", "0", "inf", 0.5, "modules/core/hyperbolic/unit/scalar/csch.cpp:40"),
    ("acoth", "inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/acoth.cpp:38"),
    ("acoth", "-inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/acoth.cpp:39"),
    ("acoth", "-1", "-inf", 0.5, "modules/core/hyperbolic/unit/scalar/acoth.cpp:40"),
    ("acoth", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/acoth.cpp:41"),
    ("acoth", "1", "inf", 0.5, "modules/core/hyperbolic/unit/scalar/acoth.cpp:42"),
    ("acoth", "0", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/acoth.cpp:43"),
    ("asech", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/asech.cpp:36"),
    ("asech", "0", "inf", 0.5, "modules/core/hyperbolic/unit/scalar/asech.cpp:37"),
    ("asech", "1", "0", 0.5, "modules/core/hyperbolic/unit/scalar/asech.cpp:39"),
    ("acsch", "inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/acsch.cpp:36"),
    ("acsch", "-inf", "0", 0.5, "modules/core/hyperbolic/unit/scalar/acsch.cpp:37"),
    ("acsch", "nan", "nan", 0.5, "modules/core/hyperbolic/unit/scalar/acsch.cpp:38"),
    ("acsch", "-0", "-inf", 0.5, "modules/core/hyperbolic/unit/scalar/acsch.cpp:39"),
    ("acsch", "0", "inf", 0.5, "modules/core/hyperbolic/unit/scalar/acsch.cpp:40"),
    ("acot", "inf", "0", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:43"),
    ("acot", "-inf", "0", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:44"),
    ("acot", "nan", "nan", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:45"),
    ("acot", "0.5", "1.107148717794090", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:47"),
    ("acot", "-0.5", "-1.107148717794090", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:48"),
    ("acot", "-1", "-7.853981633974483e-01", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:49"),
    ("acot", "1", "7.853981633974483e-01", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:50"),
    ("acot", "0", "1.570796326794897", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:51"),
    ("acot", "1/eps", "eps", 1.5, "modules/core/trigonometric/unit/scalar/acot.cpp:52"),
    ("asec", "inf", "pi/2", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:45"),
    ("asec", "-inf", "pi/2", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:46"),
    ("asec", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:47"),
    ("asec", "0", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:48"),
    ("asec", "-2", "2pi/3", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:50"),
    ("asec", "-1", "pi", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:51"),
    ("asec", "1", "0", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:52"),
    ("asec", "2", "pi/3", 0.5, "modules/core/trigonometric/unit/scalar/asec.cpp:53"),
    ("acsc", "inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:44"),
    ("acsc", "-inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:45"),
    ("acsc", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:46"),
    ("acsc", "0", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:47"),
    ("acsc", "-2", "-pi/6", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:49"),
    ("acsc", "-1", "-pi/2", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:50"),
    ("acsc", "1", "pi/2", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:51"),
    ("acsc", "2", "pi/6", 0.5, "modules/core/trigonometric/unit/scalar/acsc.cpp:52"),
    ("acotd", "inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/acotd.cpp:40"),
    ("acotd", "-inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/acotd.cpp:41"),
    ("acotd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acotd.cpp:42"),
    ("acotd", "-0", "-90", 0.5, "modules/core/trigonometric/unit/scalar/acotd.cpp:44"),
    ("acotd", "-1", "-45", 0.5, "modules/core/trigonometric/unit/scalar/acotd.cpp:45"),
    ("acotd", "1", "45", 0.5, "modules/core/trigonometric/unit/scalar/acotd.cpp:46"),
    ("acotd", "0", "90", 0.5, "modules/core/trigonometric/unit/scalar/acotd.cpp:47"),
    ("asecd", "inf", "90", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:41"),
    ("asecd", "-inf", "90", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:42"),
    ("asecd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:43"),
    ("asecd", "0", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:44"),
    ("asecd", "-2", "120", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:46"),
    ("asecd", "-1", "180", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:47"),
    ("asecd", "1", "0", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:48"),
    ("asecd", "2", "60", 0.5, "modules/core/trigonometric/unit/scalar/asecd.cpp:49"),
    ("acscd", "inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:41"),
    ("acscd", "-inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:42"),
    ("acscd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:43"),
    ("acscd", "0", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:44"),
    ("acscd", "-2", "-30", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:46"),
    ("acscd", "-1", "-90", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:47"),
    ("acscd", "1", "90", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:48"),
    ("acscd", "2", "30", 0.5, "modules/core/trigonometric/unit/scalar/acscd.cpp:49"),
    ("acosd", "inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:42"),
    ("acosd", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:43"),
    ("acosd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:44"),
    ("acosd", "0.5", "60", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:46"),
    ("acosd", "-0.5", "120", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:47"),
    ("acosd", "-1", "180", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:48"),
    ("acosd", "1", "0", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:49"),
    ("acosd", "0", "90", 0.5, "modules/core/trigonometric/unit/scalar/acosd.cpp:50"),
    ("acospi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/acospi.cpp:40"),
    ("acospi", "0.5", "1/3", 0.5, "modules/core/trigonometric/unit/scalar/acospi.cpp:42"),
    ("acospi", "-0.5", "2/3", 0.5, "modules/core/trigonometric/unit/scalar/acospi.cpp:43"),
    ("acospi", "-1", "1", 0.5, "modules/core/trigonometric/unit/scalar/acospi.cpp:44"),
    ("acospi", "1", "0", 0.5, "modules/core/trigonometric/unit/scalar/acospi.cpp:45"),
    ("acospi", "0", "0.5", 0.5, "modules/core/trigonometric/unit/scalar/acospi.cpp:46"),
    ("asind", "inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:42"),
    ("asind", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:43"),
    ("asind", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:44"),
    ("asind", "0.5", "30", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:46"),
    ("asind", "-0.5", "-30", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:47"),
    ("asind", "-1", "-90", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:48"),
    ("asind", "1", "90", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:49"),
    ("asind", "0", "0", 0.5, "modules/core/trigonometric/unit/scalar/asind.cpp:50"),
    ("asinpi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/asinpi.cpp:40"),
    ("asinpi", "0.5", "1/6", 0.5, "modules/core/trigonometric/unit/scalar/asinpi.cpp:42"),
    ("asinpi", "-0.5", "-1/6", 0.5, "modules/core/trigonometric/unit/scalar/asinpi.cpp:43"),
    ("asinpi", "-1", "-0.5", 0.5, "modules/core/trigonometric/unit/scalar/asinpi.cpp:44"),
    ("asinpi", "1", "0.5", 0.5, "modules/core/trigonometric/unit/scalar/asinpi.cpp:45"),
    ("asinpi", "0", "0", 0.5, "modules/core/trigonometric/unit/scalar/asinpi.cpp:46"),
    ("atand", "inf", "90", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:42"),
    ("atand", "-inf", "-90", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:43"),
    ("atand", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:44"),
    ("atand", "0.5", "2.656505117707799e+01", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:46"),
    ("atand", "-0.5", "-2.656505117707799e+01", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:47"),
    ("atand", "-1", "-45", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:48"),
    ("atand", "1", "45", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:49"),
    ("atand", "0", "0", 0.5, "modules/core/trigonometric/unit/scalar/atand.cpp:50"),
    ("atanpi", "inf", "0.5", 0.5, "modules/core/trigonometric/unit/scalar/atanpi.cpp:40"),
    ("atanpi", "-inf", "-0.5", 0.5, "modules/core/trigonometric/unit/scalar/atanpi.cpp:41"),
    ("atanpi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/atanpi.cpp:42"),
    ("atanpi", "-1", "-0.25", 0.5, "modules/core/trigonometric/unit/scalar/atanpi.cpp:44"),
    ("atanpi", "1", "0.25", 0.5, "modules/core/trigonometric/unit/scalar/atanpi.cpp:45"),
    ("atanpi", "0", "0", 0.5, "modules/core/trigonometric/unit/scalar/atanpi.cpp:46"),
    ("indeg", "inf", "inf", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:44"),
    ("indeg", "-inf", "-inf", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:45"),
    ("indeg", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:46"),
    ("indeg", "-pi", "-180", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:48"),
    ("indeg", "-pi/2", "-90", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:49"),
    ("indeg", "-pi/4", "-45", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:50"),
    ("indeg", "pi", "180", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:51"),
    ("indeg", "pi/2", "90", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:52"),
    ("indeg", "pi/4", "45", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:53"),
    ("indeg", "0", "0", 0.5, "modules/core/trigonometric/unit/scalar/indeg.cpp:54"),
    ("inrad", "inf", "inf", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:44"),
    ("inrad", "-inf", "-inf", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:45"),
    ("inrad", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:46"),
    ("inrad", "-180", "-pi", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:48"),
    ("inrad", "-45", "-pi/4", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:49"),
    ("inrad", "-90", "-pi/2", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:50"),
    ("inrad", "0", "0", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:51"),
    ("inrad", "180", "pi", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:52"),
    ("inrad", "45", "pi/4", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:53"),
    ("inrad", "90", "pi/2", 0.5, "modules/core/trigonometric/unit/scalar/inrad.cpp:54"),
    ("secd", "-90", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:41"),
    ("secd", "inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:42"),
    ("secd", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:43"),
    ("secd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:44"),
    ("secd", "90", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:45"),
    ("secd", "-180", "-1", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:47"),
    ("secd", "-45", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:48"),
    ("secd", "0", "1", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:49"),
    ("secd", "180", "-1", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:50"),
    ("secd", "45", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/secd.cpp:51"),
    ("cscd", "-0", "-inf", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:41"),
    ("cscd", "-180", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:42"),
    ("cscd", "inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:43"),
    ("cscd", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:44"),
    ("cscd", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:45"),
    ("cscd", "0", "inf", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:46"),
    ("cscd", "180", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:47"),
    ("cscd", "-45", "-sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:49"),
    ("cscd", "-90", "-1", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:50"),
    ("cscd", "45", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:51"),
    ("cscd", "90", "1", 0.5, "modules/core/trigonometric/unit/scalar/cscd.cpp:52"),
    ("secpi", "0.5", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:44"),
    ("secpi", "inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:45"),
    ("secpi", "-0.5", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:46"),
    ("secpi", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:47"),
    ("secpi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:48"),
    ("secpi", "-0.25", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:50"),
    ("secpi", "1", "-1", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:51"),
    ("secpi", "0.25", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:52"),
    ("secpi", "0", "1", 0.5, "modules/core/trigonometric/unit/scalar/secpi.cpp:53"),
    ("cscpi", "-0", "-inf", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:44"),
    ("cscpi", "inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:45"),
    ("cscpi", "-inf", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:46"),
    ("cscpi", "-1", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:47"),
    ("cscpi", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:48"),
    ("cscpi", "1", "nan", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:49"),
    ("cscpi", "0", "inf", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:50"),
    ("cscpi", "-0.25", "-sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:52"),
    ("cscpi", "0.5", "1", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:53"),
    ("cscpi", "-0.5", "-1", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:54"),
    ("cscpi", "0.25", "sqrt(2)", 0.5, "modules/core/trigonometric/unit/scalar/cscpi.cpp:55"),
    ("sinc", "inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:43"),
    ("sinc", "-inf", "0", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:44"),
    ("sinc", "nan", "nan", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:45"),
    ("sinc", "-pi/2", "0.6366197723675814", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:47"),
    ("sinc", "-pi/4", "0.9003163161571061", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:48"),
    ("sinc", "pi/2", "0.6366197723675814", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:49"),
    ("sinc", "pi/4", "0.9003163161571061", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:50"),
    ("sinc", "eps", "1", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:51"),
    ("sinc", "mindenormal", "1", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:52"),
    ("sinc", "0", "1", 0.5, "modules/core/trigonometric/unit/scalar/sinc.cpp:53"),
    ("sincpi", "inf", "0", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:41"),
    ("sincpi", "-inf", "0", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:42"),
    ("sincpi", "nan", "nan", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:43"),
    ("sincpi", "-0.5", "0.6366197723675814", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:45"),
    ("sincpi", "-0.25", "0.9003163161571061", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:46"),
    ("sincpi", "0.5", "0.6366197723675814", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:47"),
    ("sincpi", "0.25", "0.9003163161571061", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:48"),
    ("sincpi", "eps", "1", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:49"),
    ("sincpi", "mindenormal", "1", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:50"),
    ("sincpi", "0", "1", 1.0, "modules/core/trigonometric/unit/scalar/sincpi.cpp:51"),
]


def build_kats():
    rows = [dict(func=f, x=x, expected=e, tol_ulpdist=t, source=s) for f, x, e, t, s in KATS]
    with open(os.path.join(HERE, "kats.json"), "w") as fh:
        json.dump(rows, fh, indent=1)


if __name__ == "__main__":
    build_math(np.float32, "f32")
    build_math(np.float64, "f64")
    build_bs()
    build_bs64()
    build_kats()
    print("golden fixtures written")
