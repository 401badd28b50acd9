#!/usr/bin/env python3
"""Fits the polynomial approximations used by the accurate Black-Scholes
kernel (nt2_amd/include/nt2/hip/bs_accurate.hpp) and writes them as a header.

* erfcx(z) = exp(z^2) erfc(z) on z in [0, ZMAX]: polynomial of degree DEG in
  t = A*y + B, y = C/(C+z), t in [-1, 1] (the change of variable that
  Faddeeva-package style erfcx approximations use), least-squares on
  Chebyshev nodes with mpmath values, converted to monomial form in t.
* exp(r) on |r| <= ln(2)/2: degree-10 polynomial fitted the same way.

Usage: python3 gen_bs_coeffs.py <out.hpp>
"""
import sys

import mpmath
import numpy as np

mpmath.mp.dps = 40

ZMAX = 12.0
C = 2.5
DEG = 12
EXP_DEG = 10


def fit(f, lo, hi, deg, nodes=600):
    k = np.arange(nodes)
    x = (lo + hi) / 2 + (hi - lo) / 2 * np.cos(np.pi * (k + 0.5) / nodes)
    y = np.array([float(f(mpmath.mpf(v))) for v in x])
    # weighted for relative error
    t = (2 * x - (lo + hi)) / (hi - lo)
    V = np.polynomial.chebyshev.chebvander(t, deg) / y[:, None]
    cheb, *_ = np.linalg.lstsq(V, np.ones_like(y), rcond=None)
    mono = np.polynomial.chebyshev.cheb2poly(cheb)
    return mono  # coefficients of t^0 .. t^deg


def main():
    out = sys.argv[1]
    ymin = C / (C + ZMAX)
    # erfcx as a function of t
    def erfcx_t(t):
        y = (t * (1 - ymin) + (1 + ymin)) / 2
        z = C / y - C
        return mpmath.exp(z * z) * mpmath.erfc(z)
    ce = fit(erfcx_t, -1.0, 1.0, DEG)
    A = 2.0 / (1 - ymin)
    B = -(1 + ymin) / (1 - ymin)
    half = float(mpmath.log(2) / 2)
    cx = fit(mpmath.exp, -half, half, EXP_DEG)
    # map exp's t back to r: t = r/half -> coefficients in r
    cx = [c / half ** i for i, c in enumerate(cx)]
    # verify
    zt = np.linspace(0, ZMAX, 20001)
    yt = C / (C + zt)
    tt = A * yt + B
    approx = np.polynomial.polynomial.polyval(tt, ce)
    from scipy.special import erfcx
    e1 = np.max(np.abs(approx / erfcx(zt) - 1))
    rt = np.linspace(-half, half, 20001)
    e2 = np.max(np.abs(np.polynomial.polynomial.polyval(rt, cx) / np.exp(rt) - 1))
    with open(out, "w") as f:
        f.write("// Generated by nt2_amd/csrc/gen_bs_coeffs.py. Do not edit.\n")
        f.write("// erfcx on [0, %g]: max relative error %.3g; exp on |r| <= ln2/2: %.3g\n" % (ZMAX, e1, e2))
        f.write("#pragma once\nnamespace nt2 { namespace hip { namespace bsacc {\n")
        f.write("constexpr double kErfcxZmax = %r;\n" % ZMAX)
        f.write("constexpr double kErfcxC = %r;\n" % C)
        f.write("constexpr double kErfcxA = %r;\n" % A)
        f.write("constexpr double kErfcxB = %r;\n" % B)
        f.write("// erfcx(z) ~ sum_k kErfcx[k] t^k, highest degree first\n")
        f.write("constexpr double kErfcx[%d] = {\n" % (DEG + 1))
        for c in ce[::-1]:
            f.write("    %r,\n" % float(c))
        f.write("};\n")
        f.write("// exp(r) ~ sum_k kExp[k] r^k, highest degree first\n")
        f.write("constexpr double kExp[%d] = {\n" % (EXP_DEG + 1))
        for c in cx[::-1]:
            f.write("    %r,\n" % float(c))
        f.write("};\n}}}  // namespace nt2::hip::bsacc\n")
    print("erfcx max rel err %.3g, exp max rel err %.3g" % (e1, e2))


if __name__ == "__main__":
    main()
