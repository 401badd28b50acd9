#!/usr/bin/env python3
"""Generate the 2/pi and pi/2 digit tables used by the Payne-Hanek reduction.

The tables are mathematical constants (digits of 2/pi, chunks of pi/2 and the
high words of n*pi/2), computed here with mpmath at 700 digits and written as
C++ headers.  The same script writes the copy used by the HIP device code
(nt2_amd/include/nt2/hip/detail/pio2_tables.hpp) and the copy used by the CPU
oracle (oracle/pio2_tables.h); the layouts follow the fdlibm reduction that
NT2 embeds in modules/core/trigonometric/include/nt2/trigonometric/functions/
scalar/rem_pio2.hpp:63-117 (single) and :482-520 (double).

Usage: python3 gen_pio2_tables.py <out.hpp> <namespace> [--device]
With --device the arrays are emitted as __constant__ (HIP constant memory).
"""
import struct
import sys

import mpmath

mpmath.mp.dps = 700


def two_over_pi_chunks(bits, count):
    """First `count` chunks of `bits` bits of the fraction 2/pi (value < 1)."""
    v = mpmath.mpf(2) / mpmath.pi
    out = []
    scale = 2 ** bits
    for _ in range(count):
        v *= scale
        d = int(mpmath.floor(v))
        out.append(d)
        v -= d
    return out


def pio2_chunks(bits_first, bits, count, fmt):
    """pi/2 cut into floating chunks: first holds `bits_first` leading bits,
    each following one the next `bits` bits (fdlibm's PIo2[] tables)."""
    v = mpmath.pi / 2
    out = []
    # exponent of pi/2 is 0 (1 <= pi/2 < 2); chunk k covers bits
    # [8k, 8k+8) (single) or [24k, 24k+24) (double) below the binary point
    # counted from the leading bit.
    pos = bits_first
    chunk = mpmath.floor(v * 2 ** (pos - 1)) / 2 ** (pos - 1)
    out.append(chunk)
    rem = v - chunk
    for _ in range(count - 1):
        pos += bits
        c = mpmath.floor(rem * 2 ** (pos - 1)) / 2 ** (pos - 1)
        out.append(c)
        rem -= c
    packed = []
    for c in out:
        f = float(c)
        if fmt == "f":
            assert float(struct.unpack("<f", struct.pack("<f", f))[0]) == f
        packed.append(f)
    return packed


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def f64_bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def npio2_hw_single():
    # high bits of n*pi/2 rounded to float, last 8 bits cleared
    out = []
    for n in range(1, 33):
        v = float(mpmath.mpf(n) * mpmath.pi / 2)
        b = f32_bits(struct.unpack("<f", struct.pack("<f", v))[0])
        out.append(b & 0xFFFFFF00)
    return out


def npio2_hw_double():
    out = []
    for n in range(1, 33):
        v = float(mpmath.mpf(n) * mpmath.pi / 2)
        out.append(f64_bits(v) >> 32)
    return out


def main():
    out_path, ns = sys.argv[1], sys.argv[2]
    q = "static __constant__ const" if "--device" in sys.argv[3:] else "static const"
    t8 = two_over_pi_chunks(8, 198)
    t24 = two_over_pi_chunks(24, 66)
    p8 = pio2_chunks(8, 8, 11, "f")
    p24 = pio2_chunks(24, 24, 8, "d")
    hw32 = npio2_hw_single()
    hw64 = npio2_hw_double()

    def rows(vals, fmtf, per):
        s = []
        for i in range(0, len(vals), per):
            s.append("  " + ", ".join(fmtf(v) for v in vals[i:i + per]) + ",")
        return "\n".join(s)

    with open(out_path, "w") as f:
        f.write("// Generated by nt2_amd/csrc/gen_pio2_tables.py (mpmath, 700 digits). Do not edit.\n")
        f.write("// Digits of 2/pi, pi/2 cut in 8- and 24-bit chunks, and high words of n*pi/2.\n")
        f.write("#pragma once\n#include <stdint.h>\n")
        f.write("namespace %s {\n" % ns)
        f.write("// 2/pi in 8-bit chunks (single-precision Payne-Hanek)\n")
        f.write(q + " int32_t two_over_pi_8[%d] = {\n%s\n};\n" %
                (len(t8), rows(t8, lambda v: "0x%02X" % v, 12)))
        f.write("// 2/pi in 24-bit chunks (double-precision Payne-Hanek)\n")
        f.write(q + " int32_t two_over_pi_24[%d] = {\n%s\n};\n" %
                (len(t24), rows(t24, lambda v: "0x%06X" % v, 6)))
        f.write("// pi/2 in 8-bit float chunks\n")
        f.write(q + " uint32_t pio2_8_bits[%d] = {\n%s\n};\n" %
                (len(p8), rows([f32_bits(v) for v in p8], lambda v: "0x%08Xu" % v, 4)))
        f.write("// pi/2 in 24-bit double chunks\n")
        f.write(q + " uint64_t pio2_24_bits[%d] = {\n%s\n};\n" %
                (len(p24), rows([f64_bits(v) for v in p24], lambda v: "0x%016Xull" % v, 2)))
        f.write("// float(n*pi/2) & 0xffffff00, n = 1..32\n")
        f.write(q + " int32_t npio2_hw_f[32] = {\n%s\n};\n" %
                rows(hw32, lambda v: "0x%08X" % v, 6))
        f.write("// high word of double(n*pi/2), n = 1..32\n")
        f.write(q + " int32_t npio2_hw_d[32] = {\n%s\n};\n" %
                rows(hw64, lambda v: "0x%08X" % v, 6))
        f.write("}  // namespace %s\n" % ns)


if __name__ == "__main__":
    main()
