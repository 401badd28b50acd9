#!/usr/bin/env python3
"""Fits the tables of the accurate Black-Scholes policy
(nt2_amd/include/nt2/hip/bs_accurate.hpp) and writes them as a header.

* N(-|d|) = exp(-d^2/2) * erfcx(|d|/sqrt2) / 2.  The erfcx(.)/2 factor is a
  piecewise polynomial in the variable y = C/(C + |d|) (the change of variable
  Faddeeva-style erfcx codes use: erfcx is close to linear in y for large
  arguments).  y in [C/(C + DMAX), 1] is mapped to x = A*y + B in [0, 32);
  interval j = floor(x) holds a degree-DEG polynomial in u = x - j in [0, 1),
  fitted for relative error on Chebyshev nodes with mpmath values.  The
  coefficients are stored coefficient-major (32 doubles per power), so that a
  wave's lookups of one coefficient touch 32 distinct LDS bank pairs: no bank
  conflicts, whatever the intervals of the 64 lanes.
* exp(r) on |r| <= ln2/64 (the reduced argument of exp(w) = 2^(k/32) e^r):
  degree-3 fit for relative error.
* 2^(j/32), j = 0..31, correctly rounded.
* exp(w) on |w| <= SMALL_W (the discount factor exp(-rT) for usual rates and
  maturities): degree-SMALL_DEG fit for relative error, so that the common
  case needs no table lookup and no scaling.

Usage: python3 gen_bs_coeffs.py <out.hpp>
"""
import sys

import mpmath
import numpy as np

mpmath.mp.dps = 40

NINT = 32
DEG = 4
DMAX = 16.97                      # |d| clamp (= 12 sqrt2 - 0.0005): N(-16.97) < 1e-63
C = 2.5 * float(mpmath.sqrt(2))   # y = C / (C + |d|)
EXP_DEG = 3
SMALL_W = 0.25
SMALL_DEG = 7


def d_of_x(x, A, B):
    y = (x - B) / A
    return C / y - C


def half_erfcx_mp(d):
    z = mpmath.mpf(d) / mpmath.sqrt(2)
    return mpmath.exp(z * z) * mpmath.erfc(z) / 2


def main():
    out = sys.argv[1]
    ymin = C / (C + DMAX)
    A = 31.9999 / (1.0 - ymin)     # x(1) = 31.9999, so floor(x) <= 31 for every |d| >= 0
    B = -A * ymin                  # x(ymin) = 0
    nodes = 40
    k = np.arange(nodes)
    uc = 0.5 + 0.5 * np.cos(np.pi * (k + 0.5) / nodes)
    ut = np.linspace(0.0, 1.0, 201)
    rows = []
    worst = 0.0
    for j in range(NINT):
        fv = np.array([float(half_erfcx_mp(d_of_x(j + u, A, B))) for u in uc])
        V = np.vander(uc, DEG + 1, increasing=True) / fv[:, None]
        c, *_ = np.linalg.lstsq(V, np.ones_like(fv), rcond=None)
        ref = np.array([float(half_erfcx_mp(d_of_x(j + u, A, B))) for u in ut])
        worst = max(worst, float(np.max(np.abs(np.polynomial.polynomial.polyval(ut, c) / ref - 1))))
        rows.append(c)
    half = float(mpmath.log(2)) / 64
    xe = half * np.cos(np.pi * (np.arange(200) + 0.5) / 200)
    ev = np.array([float(mpmath.exp(mpmath.mpf(x))) for x in xe])
    V = np.vander(xe, EXP_DEG + 1, increasing=True) / ev[:, None]
    ce, *_ = np.linalg.lstsq(V, np.ones_like(xe), rcond=None)
    xt = np.linspace(-half, half, 4001)
    et = np.array([float(mpmath.exp(mpmath.mpf(x))) for x in xt])
    eworst = float(np.max(np.abs(np.polynomial.polynomial.polyval(xt, ce) / et - 1)))
    exp2 = [float(mpmath.mpf(2) ** (mpmath.mpf(j) / 32)) for j in range(32)]
    xs = SMALL_W * np.cos(np.pi * (np.arange(200) + 0.5) / 200)
    sv = np.array([float(mpmath.exp(mpmath.mpf(x))) for x in xs])
    V = np.vander(xs, SMALL_DEG + 1, increasing=True) / sv[:, None]
    cs, *_ = np.linalg.lstsq(V, np.ones_like(xs), rcond=None)
    xt = np.linspace(-SMALL_W, SMALL_W, 4001)
    st = np.array([float(mpmath.exp(mpmath.mpf(x))) for x in xt])
    sworst = float(np.max(np.abs(np.polynomial.polynomial.polyval(xt, cs) / st - 1)))
    with open(out, "w") as f:
        f.write("// Generated by nt2_amd/csrc/gen_bs_coeffs.py. Do not edit.\n")
        f.write("// erfcx(|d|/sqrt2)/2 on |d| <= %g: max relative error %.3g;\n" % (DMAX, worst))
        f.write("// exp(r) on |r| <= ln2/64: max relative error %.3g;\n" % eworst)
        f.write("// exp(w) on |w| <= %g: max relative error %.3g\n" % (SMALL_W, sworst))
        f.write("#pragma once\nnamespace nt2 { namespace hip { namespace bsacc {\n")
        f.write("constexpr int kTailN = %d;      // intervals\n" % NINT)
        f.write("constexpr int kTailDeg = %d;    // polynomial degree\n" % DEG)
        f.write("constexpr double kTailDmax = %r;\n" % DMAX)
        f.write("constexpr double kTailC = %r;   // y = C / (C + |d|)\n" % C)
        f.write("constexpr double kTailA = %r;   // x = A y + B in [0, 32)\n" % A)
        f.write("constexpr double kTailB = %r;\n" % B)
        f.write("// coefficient-major: kTail[i * 32 + j] multiplies u^i on interval j\n")
        f.write("constexpr double kTail[%d] = {\n" % (NINT * (DEG + 1)))
        for i in range(DEG + 1):
            for j in range(NINT):
                f.write("    %r,\n" % float(rows[j][i]))
        f.write("};\n")
        f.write("// exp(r) ~ sum_i kExpR[i] r^i\n")
        f.write("constexpr double kExpR[%d] = {%s};\n" % (EXP_DEG + 1, ", ".join(repr(float(v)) for v in ce)))
        f.write("// 2^(j/32), j = 0..31\n")
        f.write("constexpr double kExp2j[32] = {\n")
        for v in exp2:
            f.write("    %r,\n" % v)
        f.write("};\n")
        f.write("// exp(w) ~ sum_i kExpS[i] w^i for |w| <= kExpSmall\n")
        f.write("constexpr double kExpSmall = %r;\n" % SMALL_W)
        f.write("constexpr double kExpS[%d] = {%s};\n" % (SMALL_DEG + 1, ", ".join(repr(float(v)) for v in cs)))
        f.write("}}}  // namespace nt2::hip::bsacc\n")
    print("tail max rel err %.3g, exp max rel err %.3g, small exp max rel err %.3g" % (worst, eworst, sworst))


if __name__ == "__main__":
    main()
