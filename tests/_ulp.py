"""ULP measures used by the tests and the bench.

classic_ulp(a, ref): |a - ref| in units of the spacing of `a`'s dtype at
ref (ref given in float64 or exactly representable), NaN == NaN counts 0.
nt2_ulpdist(a, b): NT2's ulpdist, which counts half-ULPs:
modules/boost/simd/base/include/boost/simd/ieee/functions/scalar/ulpdist.hpp:66-82.
"""
import numpy as np


def classic_ulp(a, ref, dtype=None):
    a = np.asarray(a)
    dtype = dtype or a.dtype
    ref = np.asarray(ref, dtype=np.float64)
    a64 = a.astype(np.float64)
    info = np.finfo(dtype)
    mag = np.abs(ref)
    # spacing of dtype at |ref|: 2^(floor(log2|ref|) - mantissa bits), floored at the subnormal step
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.floor(np.log2(np.where(mag > 0, mag, info.tiny)))
    e = np.maximum(e, np.log2(info.tiny))
    spacing = np.exp2(e - info.nmant)
    err = np.abs(a64 - ref) / spacing
    both_nan = np.isnan(a64) & np.isnan(ref)
    same = (a64 == ref)
    err = np.where(both_nan | same, 0.0, err)
    err = np.where(np.isnan(err), np.inf, err)
    return err


def nt2_ulpdist(a0, a1):
    """Scalar NT2 ulpdist (half-ULP units)."""
    t = type(a0)
    if a0 == a1:
        return 0.0
    if np.isnan(a0) and np.isnan(a1):
        return 0.0
    m1, e1 = np.frexp(a0)
    m2, e2 = np.frexp(a1)
    expo = -max(int(e1), int(e2))
    if e1 == e2:
        e = abs(m1 - m2)
    else:
        e = abs(np.ldexp(a0, expo) - np.ldexp(a1, expo))
    eps = np.finfo(t).eps
    return float(e / eps)


def dd_ulp(a, hi, lo):
    """|a - (hi + lo)| in units of the float64 spacing at hi, for a reference
    price given as the double-double hi + lo (a - hi is exact when a is
    within a few ulp of hi)."""
    a = np.asarray(a, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    lo = np.asarray(lo, dtype=np.float64)
    info = np.finfo(np.float64)
    mag = np.abs(hi)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.floor(np.log2(np.where(mag > 0, mag, info.tiny)))
    e = np.maximum(e, np.log2(info.tiny))
    spacing = np.exp2(e - info.nmant)
    err = np.abs((a - hi) - lo) / spacing
    return np.where(np.isnan(err), np.inf, err)
