"""The image demos on the GPU through the C ABI (nt2h_rgb2grey_f32,
nt2h_rgb2yuv_f32, nt2h_linspace_*, nt2h_mandelbrot_f32) against the oracle:
bit-identical under the exact policy, ragged sizes and unaligned planes
included; the fast policy within one rounding per fused operation."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import _oracle as O  # noqa: E402
from _gpu import first_diff, same_bits, say  # noqa: E402
from nt2_amd import EXACT, FAST, ops  # noqa: E402


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("n,offset", [(1, 0), (7, 0), (1000, 0), (1000, 1), (128 * 128, 0),
                                      (1 << 20, 0), ((1 << 20) + 3, 3)])
def test_rgb2grey_rgb2yuv_exact(n, offset):
    rng = np.random.default_rng(n + offset)
    planes = [rng.uniform(0, 255, n + offset).astype(np.float32) for _ in range(3)]
    say("rgb2grey/rgb2yuv exact n=%d offset=%d" % (n, offset))
    d = [_dev(p)[offset:] for p in planes]  # an offset of 1 float breaks 16-byte alignment
    r, g, b = (p[offset:] for p in planes)
    got = ops.rgb2grey(*d, policy=EXACT).cpu().numpy()
    want = O.rgb2grey(r, g, b)
    assert same_bits(got, want), first_diff(got, want)
    y, u, v = (t.cpu().numpy() for t in ops.rgb2yuv(*d, policy=EXACT))
    wy, wu, wv = O.rgb2yuv(r, g, b)
    assert same_bits(y, wy), first_diff(y, wy)
    assert same_bits(u, wu), first_diff(u, wu)
    assert same_bits(v, wv), first_diff(v, wv)


def test_rgb2grey_demo_ramp_and_fast():
    # the demo's own input: r = g = b = i - 1 on a 128 x 128 image
    ramp = np.arange(128 * 128, dtype=np.float32)
    d = _dev(ramp)
    assert same_bits(ops.rgb2grey(d, d, d).cpu().numpy(), O.rgb2grey(ramp, ramp, ramp))
    rng = np.random.default_rng(5)
    r, g, b = (rng.uniform(0, 255, 100000).astype(np.float32) for _ in range(3))
    fast = ops.rgb2grey(_dev(r), _dev(g), _dev(b), policy=FAST).cpu().numpy().astype(np.float64)
    ex = 0.3 * r.astype(np.float64) + 0.59 * g.astype(np.float64) + 0.11 * b.astype(np.float64)
    # the float32 constants differ from the decimals by < 2^-24 relative
    assert np.all(np.abs(fast - ex) <= 3 * np.spacing(np.float32(ex)).astype(np.float64) + 1e-7 * ex)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n,lo,hi", [(1, 0.0, 9.0), (2, -1.0, 1.0), (7, 0.0, 1.0), (7, 1.0, 0.0), (257, -1.0, 1.0),
                                     (1600, -1.5, 1.5), (1000003, -3.0, 7.0)])
def test_linspace_exact(dt, n, lo, hi):
    got = ops.linspace(lo, hi, n, dtype=torch.float32 if dt is np.float32 else torch.float64).cpu().numpy()
    assert same_bits(got, O.linspace(lo, hi, n, dt))


def _grid(h, w, a0=-1.5, a1=1.5, b0=-1.5, b1=1.5):
    """The demo's A = expand_to(linspace(a0, a1, w), h, w) and
    B = expand_to(colvect(linspace(b0, b1, h)), h, w), flat column-major."""
    la, lb = O.linspace(a0, a1, w), O.linspace(b0, b1, h)
    A = np.repeat(la, h)             # A(i, j) = la(j)
    B = np.tile(lb, w)               # B(i, j) = lb(i)
    return A.astype(np.float32), B.astype(np.float32)


@pytest.mark.parametrize("h,w", [(100, 100), (400, 400), (1600, 1600), (37, 53), (1, 13)])
def test_mandelbrot_exact(h, w):
    A, B = _grid(h, w)
    say("mandelbrot exact %dx%d" % (h, w))
    got = ops.mandelbrot(_dev(A), _dev(B)).cpu().numpy()
    want = O.mandelbrot(A, B)
    assert np.array_equal(got, want), first_diff(got, want)


def test_mandelbrot_packs_tails_and_limits():
    rng = np.random.default_rng(9)
    n = 8 * 1000 + 5
    a = rng.uniform(-2.2, 1.2, n).astype(np.float32)
    b = rng.uniform(-1.6, 1.6, n).astype(np.float32)
    a[::8] = 0.0  # lane 0 of each pack never escapes
    b[::8] = 0.0
    for it in (1, 2, 17, 256, 1000):
        got = ops.mandelbrot(_dev(a), _dev(b), max_iter=it).cpu().numpy()
        want = O.mandelbrot(a, b, max_iter=it)
        assert np.array_equal(got, want), (it, first_diff(got, want))
    from nt2_amd.native import Nt2hError
    with pytest.raises(Nt2hError):
        ops.mandelbrot(_dev(a), _dev(b), max_iter=0)


def test_mandelbrot_fast_close_to_exact():
    A, B = _grid(800, 800)
    ex = ops.mandelbrot(_dev(A), _dev(B), policy=EXACT).cpu().numpy()
    fa = ops.mandelbrot(_dev(A), _dev(B), policy=FAST).cpu().numpy()
    # fused steps move orbits near the boundary; the set's interior and the
    # far outside do not change
    same = np.mean(ex == fa)
    say("mandelbrot fast == exact on %.4f of the pixels" % same)
    assert same > 0.95
    assert np.mean(fa[ex <= 2] == ex[ex <= 2]) > 0.999
    assert np.mean(fa[ex == 256] == 256) > 0.99


@pytest.mark.parametrize("offset", [0, 5])
def test_rgb2grey_u8_every_colour(offset):
    """All 2^24 colours as one image (plus a misaligned start): equal to the
    oracle byte for byte."""
    v = np.arange(256, dtype=np.uint8)
    planes = [a.ravel() for a in np.meshgrid(v, v, v, indexing="ij")]
    planes = [np.concatenate([np.zeros(offset, np.uint8), p, v[:11]]) for p in planes]  # ragged tail
    say("rgb2grey_u8 offset=%d" % offset)
    d = [_dev(p)[offset:] for p in planes]
    got = ops.rgb2grey_u8(*d).cpu().numpy()
    want = O.rgb2grey_u8(*(p[offset:] for p in planes))
    assert np.array_equal(got, want), first_diff(got, want)
