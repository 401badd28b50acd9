"""Pins the oracle's image demos (CPU only): rgb2grey, rgb2yuv, linspace and
mandelbrot.

* rgb2grey / rgb2yuv: exact rational arithmetic bounds each output by the
  roundings the demo's left-to-right expression performs, the demo's own
  inputs (r = g = b = i - 1) give back i - 1 within those roundings, and numpy
  float32 vector arithmetic (one rounding per operation) agrees bit for bit.
* linspace: the known answers of NT2's unit test
  (modules/core/generative/unit/table/linspace.cpp).
* mandelbrot: escape counts fixed by exact arithmetic for orbits whose values
  are exact in float32, the interior of the main cardioid and the period-2
  disc (256), outside the radius-2 disc (1), and NT2's pack semantics: a lane
  keeps iterating while a pack-mate is inside, but an escaped orbit never
  comes back (|z| >= max(2, |c|) makes |z^2 + c| >= |z|(|z| - 1) > |z|, and
  the first escape of a c with |c| > 2 is z1 = c itself), so every lane's
  count is its first-escape time whatever its pack-mates do.
"""
from fractions import Fraction

import numpy as np
import pytest

import _oracle as O

F32 = np.float32


def _exact(v):
    return Fraction(float(v))


def test_rgb2grey_demo_inputs_and_bounds():
    # the demo fills r = g = b = gr = i - 1 (demo/rgb2grey/nt2/rgb2grey_nt2.cpp:38-39)
    n = 128 * 128
    ramp = np.arange(n, dtype=F32)
    gr = O.rgb2grey(ramp, ramp, ramp)
    # 0.3 + 0.59 + 0.11 = 1 in exact decimals; in float32 the three constants
    # and four roundings stay within a few ulp of i - 1
    assert np.all(np.abs(gr.astype(np.float64) - ramp) <= 4 * np.spacing(np.maximum(ramp, 1)))
    rng = np.random.default_rng(41)
    r, g, b = (rng.uniform(0, 255, 5000).astype(F32) for _ in range(3))
    got = O.rgb2grey(r, g, b)
    # numpy float32: one rounding per operation, in the demo's order
    want = (F32(0.3) * r + F32(0.59) * g) + F32(0.11) * b
    assert np.array_equal(got, want)
    for i in range(0, 5000, 97):
        ex = _exact(F32(0.3)) * _exact(r[i]) + _exact(F32(0.59)) * _exact(g[i]) + _exact(F32(0.11)) * _exact(b[i])
        # four roundings of quantities at most |ex|
        assert abs(_exact(got[i]) - ex) <= 2 * Fraction(float(np.spacing(F32(abs(float(ex)))))) * 1


def test_rgb2yuv_uses_the_stored_y():
    rng = np.random.default_rng(42)
    r, g, b = (rng.uniform(0, 255, 5000).astype(F32) for _ in range(3))
    y, u, v = O.rgb2yuv(r, g, b)
    yy = (F32(0.299) * r + F32(0.587) * g) + F32(0.114) * b
    assert np.array_equal(y, yy)
    assert np.array_equal(u, F32(0.492) * (b - yy))
    assert np.array_equal(v, F32(0.877) * (r - yy))
    # grey input: y = the grey level within a few ulp, u = v = small
    ramp = np.arange(4096, dtype=F32)
    y, u, v = O.rgb2yuv(ramp, ramp, ramp)
    assert np.all(np.abs(y.astype(np.float64) - ramp) <= 4 * np.spacing(np.maximum(ramp, 1)))
    assert np.all(np.abs(u) <= 4 * np.spacing(np.maximum(ramp, 1))) and np.all(np.abs(v) <= 8 * np.spacing(np.maximum(ramp, 1)))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_linspace_nt2_unit_kats(dt):
    # linspace.cpp:81-103: 7 points within 1 ulp of (1/6)(i - 1), and reversed
    x = O.linspace(0.0, 1.0, 7, dt)
    ref = np.array([dt(0) + (dt(1) / dt(6)) * dt(i) for i in range(7)], dtype=dt)
    assert np.all(np.abs(x - ref) <= np.spacing(np.maximum(np.abs(ref), dt(1e-30))))
    xr = O.linspace(1.0, 0.0, 7, dt)
    ref = np.array([dt(1) + (dt(-1) / dt(6)) * dt(i) for i in range(7)], dtype=dt)
    assert np.all(np.abs(xr - ref) <= np.spacing(np.maximum(np.abs(ref), dt(1e-30))))
    assert list(O.linspace(0.0, 9.0, 1, dt)) == [9]  # :105-107
    # :123-131: linspace(-1, 1, 257) is exactly -1 + (i - 1)/128
    assert np.array_equal(O.linspace(-1.0, 1.0, 257, dt), np.array([-1 + i / 128 for i in range(257)], dtype=dt))
    # :56-78: the default 100 points of [0, 1] within 15 ulp of (1/99)(i - 1)
    x = O.linspace(0.0, 1.0, 100, dt)
    ref = (dt(1) / dt(99)) * np.arange(100, dtype=dt)
    assert np.all(np.abs(x - ref) <= 15 * np.spacing(np.maximum(ref, dt(1e-30))))
    # :111-121: no zero between -100 and the predecessor of 0
    for k in range(8):
        z = O.linspace(-100.0, -float(np.finfo(dt).smallest_subnormal), 100 + k, dt)
        assert np.all(z != 0)
    # first and last points are the bounds
    z = O.linspace(-1.5, 1.5, 1601, dt)
    assert z[0] == dt(-1.5) and z[-1] == dt(1.5)


def _alone(a, b):
    """Each point in a pack of 8 copies of itself."""
    a, b = np.asarray(a, F32), np.asarray(b, F32)
    return O.mandelbrot(np.repeat(a, 8), np.repeat(b, 8))[::8]


def test_mandelbrot_exact_orbits():
    # c = 0, 0.25 (fixed point 1/2), -1 (cycle 0, -1), i (cycle -1+i, -i): bounded, 256
    # c = 1: |z|^2 = 0, 1, 4 -> 2;  c = -2: 0, 4 -> 1;  c = 2: 0, 4 -> 1
    # c = 0.5: |z|^2 = 0, 1/4, 9/16, 1.12890625, 2.6533..., 9.94... -> 5 (all exact in float32)
    a = [0.0, 0.25, -1.0, 0.0, 1.0, -2.0, 2.0, 0.5, 3.0]
    b = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    assert list(_alone(a, b)) == [256, 256, 256, 256, 2, 1, 1, 5, 1]
    assert list(O.mandelbrot(np.float32([0.5]), np.float32([0.0]), max_iter=3)) == [3]


def test_mandelbrot_interior_regions():
    rng = np.random.default_rng(43)
    # main cardioid: c = (1 - (1 - e^{it})^2)/4 scaled by 0.9 towards its centre
    t = rng.uniform(0, 2 * np.pi, 800)
    w = 0.9 * np.exp(1j * t) / 2
    c = w - w * w
    assert np.all(_alone(c.real, c.imag) == 256)
    # period-2 disc |c + 1| < 1/4 (scaled by 0.9)
    c2 = -1 + 0.225 * np.exp(1j * t) * rng.uniform(0, 1, 800)
    assert np.all(_alone(c2.real, c2.imag) == 256)
    # outside the radius-2 disc: only the first iteration (z = 0) counts
    c3 = 2.05 * np.exp(1j * t) * rng.uniform(1, 3, 800)
    assert np.all(_alone(c3.real, c3.imag) == 1)


def _first_escape(a, b, max_iter=256):
    x = y = F32(0)
    for k in range(max_iter):
        x2, y2 = x * x, y * y
        if not (x2 + y2 < F32(4)):
            return k
        x, y = (x2 - y2) + a, (F32(2) * x) * y + b
    return max_iter


def test_mandelbrot_pack_semantics():
    """Alone, a lane counts up to its first escape.  Next to a lane that never
    escapes it keeps iterating, and still counts nothing more."""
    g = O.linspace(-1.5, 1.5, 400)
    A, B = np.meshgrid(g, g)
    a, b = A.ravel(), B.ravel()
    alone = _alone(a, b)
    sample = np.random.default_rng(44).choice(a.size, 300, replace=False)
    for i in sample:
        assert alone[i] == _first_escape(a[i], b[i])
    # each point as lane 3 of a pack whose other lanes are c = 0
    pa = np.zeros((a.size, 8), F32)
    pb = np.zeros((a.size, 8), F32)
    pa[:, 3], pb[:, 3] = a, b
    mixed = O.mandelbrot(pa.ravel(), pb.ravel()).reshape(-1, 8)
    assert np.all(mixed[:, [0, 1, 2, 4, 5, 6, 7]] == 256)
    assert np.array_equal(mixed[:, 3], alone)
    # a tail lane is a pack of one
    n = 8 * 5 + 3
    assert np.array_equal(O.mandelbrot(a[:n], b[:n])[40:], _alone(a[40:n], b[40:n]))


def test_rgb2grey_u8_every_colour():
    """All 2^24 colours: within one grey level of the float formula (the 8.8
    weights differ from 0.3, 0.59, 0.11 by at most 0.2/256 each), grey input
    returns itself, the primaries give their weights."""
    v = np.arange(256, dtype=np.uint8)
    r, g, b = (a.ravel() for a in np.meshgrid(v, v, v, indexing="ij"))
    got = O.rgb2grey_u8(r, g, b).astype(np.int64)
    f = 0.3 * r + 0.59 * g + 0.11 * b
    assert np.all(np.abs(got - f) <= 1.0)
    assert np.array_equal(O.rgb2grey_u8(v, v, v), v)
    z = np.zeros(1, np.uint8)
    full = np.full(1, 255, np.uint8)
    assert O.rgb2grey_u8(full, z, z)[0] == 77 and O.rgb2grey_u8(z, full, z)[0] == 150
    assert O.rgb2grey_u8(z, z, full)[0] == 28
