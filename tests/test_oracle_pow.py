"""The oracle's pow (Cephes' algorithm as NT2 states it:
modules/core/exponential/include/nt2/exponential/functions/simd/common/pow_abs.hpp:65-133,
scalar/pow_abs.hpp:51-111, details/pow_kernel.hpp) pinned by NT2's own unit-test
answers, exact powers of two, float64 / mpmath values and the special-value
rules of both forms."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import _oracle as O  # noqa: E402
from _ulp import classic_ulp  # noqa: E402


def _packs(a, b, dt):
    """each (a, b) pair in a pack of identical lanes, so the pack form runs"""
    w = 8 if dt is np.float32 else 4
    a = np.repeat(np.asarray(a, dt), w)
    b = np.repeat(np.asarray(b, dt), w)
    return O.pow(a, b)[::w]


def _tail(a, b, dt):
    """each pair alone in a one-element table: NT2's scalar form"""
    return np.array([O.pow(np.array([x], dt), np.array([y], dt))[0] for x, y in zip(a, b)], dtype=dt)


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.all((a == b) & (np.signbit(a) == np.signbit(b)) | (np.isnan(a) & np.isnan(b)))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pow_nt2_unit_kats(dt):
    """modules/core/exponential/unit/simd/pow.cpp:48-58 and :148 (tolerance 0)."""
    inf, nan = np.inf, np.nan
    kats = [(inf, inf, inf), (-inf, -inf, nan), (inf, -inf, 0.0), (nan, nan, nan),
            (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0), (-1.0, 5.0, -1.0),
            (-1.0, 6.0, 1.0), (0.0, 1.0, 0.0), (2.0, -3.0, 0.125)]
    a, b, want = (np.array(c, dt) for c in zip(*kats))
    assert _same(_packs(a, b, dt), want)
    assert _same(_tail(a, b, dt), want)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pow_exact_powers_of_two(dt):
    """2^k is exact down to NT2's Powlowlim (constants/powlowlim.hpp:29-30:
    -2400 and -17183 sixteenths), so the log2 table, the sixteenths split and
    ldexp's subnormal step all have to be right; below it the result is 0."""
    lo, hi = (-149, 127) if dt is np.float32 else (-1073, 1023)
    k = np.arange(lo, hi + 1).astype(dt)
    two = np.full(k.size, 2, dt)
    assert np.array_equal(O.pow(two, k), np.exp2(k.astype(np.float64)).astype(dt))
    assert np.array_equal(O.pow(-two, k), np.where(k % 2 == 1, -1, 1) * np.exp2(k.astype(np.float64)).astype(dt))
    assert O.pow(np.array([2], dt), np.array([lo - 1], dt))[0] == 0
    assert O.pow(np.array([2], dt), np.array([hi + 1], dt))[0] == np.inf
    q = np.full(k.size, 0.5, dt)  # (1/2)^k = 2^-k
    keep = (-k >= lo) & (-k <= hi)
    assert np.array_equal(O.pow(q, k)[keep], np.exp2(-k[keep].astype(np.float64)).astype(dt))


def test_pow_f32_within_two_ulp():
    """against the float64 power (|error| of the float64 value well below
    2^-29 relative): at most 2 ulp (measured 1.83), 77 % exact."""
    rng = np.random.default_rng(5)
    n = 200000
    a = np.exp2(rng.uniform(-20, 20, n)).astype(np.float32)
    b = rng.uniform(-6, 6, n).astype(np.float32)
    got = O.pow(a, b)
    ref = np.power(a.astype(np.float64), b.astype(np.float64))
    ok = (ref >= np.finfo(np.float32).tiny) & (ref < 3e38)
    assert ok.mean() > 0.95
    err = classic_ulp(got[ok], ref[ok], np.float32)
    assert err.max() <= 2.0, err.max()
    assert 0.74 < np.mean(got[ok] == ref[ok].astype(np.float32)) < 0.80


def test_pow_f64_against_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    rng = np.random.default_rng(6)
    m = 2000
    a = np.exp2(rng.uniform(-30, 30, m))
    b = rng.uniform(-10, 10, m)
    got = O.pow(a, b)
    ref = np.array([float(mpmath.power(mpmath.mpf(x), mpmath.mpf(y))) for x, y in zip(a, b)])
    assert classic_ulp(got, ref, np.float64).max() <= 2.0


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pow_negative_bases_and_forms(dt):
    """A negative base takes the sign of (-1)^a1 for integral a1 and gives NaN
    otherwise (simd/common/pow.hpp:40-45); is_odd is is_flint((a1 - 1) / 2)
    (predicates/functions/simd/common/is_odd.hpp:47), so in f32 every a1
    above 2^24 counts as odd (a1 - 1 rounds to an even float, whose half is
    integral), while 2^24 itself is even.  Away from the
    special values both forms compute the same bits."""
    rng = np.random.default_rng(7)
    a = np.exp2(rng.uniform(-8, 8, 3000)).astype(dt)
    k = rng.integers(-20, 21, 3000).astype(dt)
    pos = O.pow(a, k)
    neg = O.pow(-a, k)
    assert _same(neg, np.where(k % 2 == 1, -pos, pos))
    frac = k + dt(0.5)
    assert np.all(np.isnan(O.pow(-a, frac)))
    if dt is np.float32:
        big = np.array([16777218.0, 16777216.0, 16777220.0, 1e30], dt)
        assert _same(_packs(-np.ones(4, dt), big, dt), np.array([-1.0, 1.0, -1.0, -1.0], dt))
    b = rng.uniform(-5, 5, 400).astype(dt)
    x = np.exp2(rng.uniform(-5, 5, 400)).astype(dt)
    assert _same(_tail(x, b, dt), _packs(x, b, dt))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pow_special_values_of_both_forms(dt):
    """The pack form replaces the computed value in the order of
    simd/common/pow_abs.hpp:117-126; the scalar form returns early in the
    order of scalar/pow_abs.hpp:56-66.  They differ only where a NaN meets an
    infinity: pow(NaN, +Inf) is 0 in packs and Inf in the tail, pow(+Inf, NaN)
    is 1 in packs and Inf in the tail (pow(-Inf, NaN) is NaN in both: a
    negative base with a non-integral exponent)."""
    v = np.array([0.0, -0.0, 0.5, 1.0, -1.0, 2.0, -2.0, 3.0, np.inf, -np.inf, np.nan], dt)
    a, b = (g.ravel() for g in np.meshgrid(v, v, indexing="ij"))
    p, t = _packs(a, b, dt), _tail(a, b, dt)
    differ = ~((p == t) | (np.isnan(p) & np.isnan(t)))
    assert differ.sum() == 2, list(zip(a[differ], b[differ]))
    i = np.flatnonzero(np.isnan(a) & (b == np.inf))[0]
    j = np.flatnonzero((a == np.inf) & np.isnan(b))[0]
    assert differ[i] and differ[j]
    assert p[i] == 0 and t[i] == np.inf and p[j] == 1 and t[j] == np.inf
    for x, y, want in [(0.0, -1.0, 0.0), (0.0, 0.0, 1.0), (0.5, np.inf, 0.0), (2.0, np.inf, np.inf),
                       (0.5, -np.inf, np.inf), (2.0, -np.inf, 0.0), (np.inf, -2.0, 0.0), (np.inf, 0.0, 1.0),
                       (-np.inf, 3.0, -np.inf), (-np.inf, 2.0, np.inf), (1.0, np.nan, 1.0),
                       (np.nan, 0.0, np.nan), (-2.0, 0.5, np.nan)]:
        assert _same(_packs([x], [y], dt), [want]), (x, y)


def test_pow_f32_error_grows_with_the_exponent():
    """NT2's f32 pow forms log2 of the significand in float
    (details/pow_kernel.hpp:32-40), so its error is multiplied by |b|: within
    2 + |b|/10 ulp of the float64 power (measured: 1 ulp below |b| = 2, 1.8 up
    to 16, 11 at |b| ~ 120)."""
    rng = np.random.default_rng(3)
    n = 300000
    a = np.exp2(rng.uniform(-4, 4, n)).astype(np.float32)
    b = rng.uniform(-120, 120, n).astype(np.float32)
    got = O.pow(a, b)
    ref = np.power(a.astype(np.float64), b.astype(np.float64))
    ok = (ref >= np.finfo(np.float32).tiny) & (ref < 3e38)
    err = classic_ulp(got[ok], ref[ok], np.float32)
    assert np.all(err <= 2.0 + np.abs(b[ok]) / 10.0)
    assert err[np.abs(b[ok]) < 2].max() <= 1.0


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pow_integral_exponent_nt2_unit_kats(dt):
    """modules/core/exponential/unit/simd/pow.cpp:77-116 (tolerance 0), NT2's
    repeated squaring for an integral exponent; and against the exact power
    wherever that is representable (small integer bases, |k| <= 12)."""
    inf, nan = np.inf, np.nan
    kats = [(inf, 3, inf), (-inf, 3, -inf), (nan, 3, nan), (inf, 4, inf), (-inf, 4, inf), (nan, 4, nan),
            (2.0, -3, 0.125), (-1.0, 3, -1.0), (1.0, 3, 1.0), (2.0, 3, 8.0), (0.0, 0, 1.0), (0.0, 3, 0.0),
            (-1.0, 4, 1.0), (1.0, 4, 1.0), (0.0, 4, 0.0)]
    a = np.array([x for x, _, _ in kats], dt)
    k = np.array([e for _, e, _ in kats])
    want = np.array([w for _, _, w in kats], dt)
    assert _same(O.pow_int(a, k), want)
    base = np.repeat(np.arange(-7, 8), 13).astype(dt)
    kk = np.tile(np.arange(0, 13), 15)
    exact = np.array([float(int(x) ** int(e)) for x, e in zip(base, kk)])
    rep = np.abs(exact) < 2.0 ** (24 if dt is np.float32 else 53)
    assert np.array_equal(O.pow_int(base, kk)[rep], exact[rep].astype(dt))
    # a negative exponent is the reciprocal of the positive power
    x = np.exp2(np.random.default_rng(2).uniform(-3, 3, 500)).astype(dt)
    assert np.array_equal(O.pow_int(x, -5), dt(1) / O.pow_int(x, 5))
