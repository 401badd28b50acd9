#!/usr/bin/env python3
"""Fits the tables of the accurate double-precision Black-Scholes policy
(nt2_amd/include/nt2/hip/bs_accurate64.hpp) and writes them as a header.

* The normal tail Q(x) = N(-x) = phi(x) R(x) for x >= 0, with R the Mills
  ratio, is computed as exp(-x^2/2) * gt(y) * y with y = C/(C + x) and
  gt(y) = R(x) (C + x) / (C sqrt(2 pi)): gt is smooth on y in (0, 1] and
  tends to 1/(C sqrt(2 pi)) as x -> infinity.  y in (0, 1] is cut into N
  pieces of width 1/N; piece j holds a degree-DEG polynomial in
  u = N y - (j + 1/2) in [-1/2, 1/2], interpolated at Chebyshev nodes with
  mpmath at 50 digits.  The constant and linear coefficients are stored as
  double-double (hi, lo) pairs, the others as doubles, coefficient-major
  (N doubles per row) so that a wave's gathers of one row touch N distinct
  LDS bank pairs.
* 2^(j/32), j = 0..31, as double-double pairs.
* ln2/32 split in three parts, the first two with 36 significant bits, so
  that k * part is exact for |k| < 2^17.

The generator prints the largest relative error of the rounded tables
against mpmath on 201 points per piece.

Usage: python3 gen_bs64_coeffs.py <out.hpp>
"""
import sys

import mpmath as mp

mp.mp.dps = 50

N = 32
DEG = 9
C = 3


def gt_of_y(y):
    y = mp.mpf(y)
    if y == 0:
        return 1 / (C * mp.sqrt(2 * mp.pi))
    x = C * (1 - y) / y
    R = mp.sqrt(mp.pi / 2) * mp.exp(x * x / 2) * mp.erfc(x / mp.sqrt(2))
    return R * (C + x) / (C * mp.sqrt(2 * mp.pi))


def split(v, bits=53):
    """v as a sum of two doubles, the first rounded to `bits` significant bits."""
    if bits == 53:
        hi = float(v)
    else:
        e = mp.floor(mp.log(abs(v), 2))
        q = mp.mpf(2) ** (e - bits + 1)
        hi = float(mp.nint(v / q) * q)
    return hi, float(v - mp.mpf(hi))


def main():
    out = sys.argv[1]
    rows = [[None] * N for _ in range(DEG + 3)]  # c0hi, c0lo, c1hi, c1lo, c2..cDEG
    worst = mp.mpf(0)
    for j in range(N):
        yc = (mp.mpf(j) + mp.mpf(1) / 2) / N
        nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / (DEG + 1)) / 2 for k in range(DEG + 1)]
        vals = [gt_of_y(yc + u / N) for u in nodes]
        A = mp.matrix([[u ** i for i in range(DEG + 1)] for u in nodes])
        c = mp.lu_solve(A, mp.matrix(vals))
        c0h, c0l = split(c[0])
        c1h, c1l = split(c[1])
        cd = [float(c[i]) for i in range(2, DEG + 1)]
        rows[0][j], rows[1][j], rows[2][j], rows[3][j] = c0h, c0l, c1h, c1l
        for i, v in enumerate(cd):
            rows[4 + i][j] = v
        for k in range(201):
            u = mp.mpf(k) / 200 - mp.mpf(1) / 2
            y = yc + u / N
            if y <= 0:
                continue
            p = mp.mpf(c0h) + mp.mpf(c0l) + (mp.mpf(c1h) + mp.mpf(c1l)) * u
            p += sum(mp.mpf(cd[i]) * u ** (i + 2) for i in range(len(cd)))
            worst = max(worst, abs(p / gt_of_y(y) - 1))
    exp2 = [split(mp.mpf(2) ** (mp.mpf(j) / 32)) for j in range(32)]
    L = mp.log(2) / 32
    L1, _ = split(L, 36)
    L2, _ = split(L - mp.mpf(L1), 36)
    L3 = float(L - mp.mpf(L1) - mp.mpf(L2))
    with open(out, "w") as f:
        f.write("// Generated by nt2_amd/csrc/gen_bs64_coeffs.py. Do not edit.\n")
        f.write("// gt(y) = R(x) (C + x) / (C sqrt(2 pi)), y = C/(C + x): max relative error %.3g\n"
                % float(worst))
        f.write("#pragma once\nnamespace nt2 { namespace hip { namespace bsacc64 {\n")
        f.write("constexpr int kGN = %d;      // pieces of y in (0, 1]\n" % N)
        f.write("constexpr int kGDeg = %d;    // polynomial degree\n" % DEG)
        f.write("constexpr int kGRows = %d;   // c0 hi, c0 lo, c1 hi, c1 lo, c2 .. c%d\n" % (DEG + 3, DEG))
        f.write("constexpr double kGC = %r;\n" % float(C))
        f.write("// coefficient-major: kG[row * %d + j] belongs to piece j\n" % N)
        f.write("constexpr double kG[%d] = {\n" % ((DEG + 3) * N))
        for r in rows:
            for v in r:
                f.write("    %r,\n" % v)
        f.write("};\n")
        f.write("// 2^(j/32) = kExp2Hi[j] + kExp2Lo[j]\n")
        f.write("constexpr double kExp2Hi[32] = {%s};\n" % ", ".join(repr(h) for h, _ in exp2))
        f.write("constexpr double kExp2Lo[32] = {%s};\n" % ", ".join(repr(l) for _, l in exp2))
        f.write("// ln2/32 = kL1 + kL2 + kL3 (kL1, kL2: 36 significant bits)\n")
        f.write("constexpr double kL1 = %r, kL2 = %r, kL3 = %r;\n" % (L1, L2, L3))
        f.write("constexpr double k32oLn2 = %r;\n" % float(32 / mp.log(2)))
        f.write("}}}  // namespace nt2::hip::bsacc64\n")
    print("gt max rel err %.3g (2^%.1f)" % (float(worst), float(mp.log(worst, 2))))


if __name__ == "__main__":
    main()
