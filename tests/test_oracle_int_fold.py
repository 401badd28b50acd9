"""Pinning the oracle's integer, fold and Black-Scholes functions (CPU only).

* uint8 ops: exhaustive over all 256 x 256 operand pairs against their closed
  forms, plus NT2's own unit-test known answers;
* sigma-delta: against an independent numpy restatement of the demo's frame
  rule (demo/sigmadelta/nt2/sigmadelta_nt2.cpp:37-58), the demo's own input
  sequence, and closed-form trajectories;
* sum / dot / saxpy: NT2's reduction KATs and exact-arithmetic bounds;
* Black-Scholes: the FP64 reference and the NT2 f32/f64 evaluations against
  the mpmath golden values (tests/golden/bs_golden.npz).
"""
import math
import os

import numpy as np
import pytest

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
A, B = np.meshgrid(np.arange(256, dtype=np.int64), np.arange(256, dtype=np.int64), indexing="ij")
A8, B8 = A.astype(np.uint8).ravel(), B.astype(np.uint8).ravel()
A_, B_ = A.ravel(), B.ravel()

CLOSED = {
    0: np.minimum(255, A_ + B_),                 # adds
    1: np.maximum(0, A_ - B_),                   # subs
    2: np.where(A_ < B_, A_ + 1, A_),            # selinc(a < b, a)
    3: np.where(A_ > B_, A_ - 1, A_),            # seldec(a > b, a)
    4: np.where(A_ < B_, 0, 1),                  # if_zero_else_one(a < b)
    5: np.abs(A_ - B_),                          # max(a,b) - min(a,b)
}


@pytest.mark.parametrize("op", sorted(CLOSED))
def test_u8_ops_exhaustive(op):
    got = O.u8_binop(op, A8, B8).astype(np.int64)
    assert np.array_equal(got, CLOSED[op])


def test_u8_kats():
    # modules/boost/simd/base/unit/arithmetic/scalar/adds.cpp: adds(Valmax, One) == Valmax,
    # adds(Zero, Zero) == Zero; subs saturates at Valmin (= 0 for uint8)
    a = np.array([255, 0, 254, 0], np.uint8)
    b = np.array([1, 0, 1, 1], np.uint8)
    assert list(O.u8_binop(0, a, b)) == [255, 0, 255, 1]
    assert list(O.u8_binop(1, a, b)) == [254, 0, 253, 0]


def sigmadelta_numpy(frames, N):
    """The demo's statements, vectorised, in int64 arithmetic."""
    fr = frames.astype(np.int64)
    bkg = fr[0].copy()
    var = np.ones_like(bkg)
    label = np.zeros_like(bkg)
    counts = [0]
    for k in range(1, fr.shape[0]):
        f = fr[k]
        b = np.where(bkg > f, bkg - 1, bkg)          # seldec(bkg > fr, bkg)
        bkg = np.where(bkg < f, b + 1, b)            # selinc(bkg < fr, ...)
        diff = np.maximum(bkg, f) - np.minimum(bkg, f)
        mul = diff.copy()
        for _ in range(N - 1):
            mul = np.minimum(255, mul + diff)        # adds
        v2 = np.where(var > mul, var - 1, var)
        v2 = np.where(var < mul, v2 + 1, v2)
        var = np.where(diff != 0, v2, var)
        label = np.where(diff < var, 0, 1)
        counts.append(int(label.sum()))
    return bkg.astype(np.uint8), var.astype(np.uint8), label.astype(np.uint8), np.array(counts)


@pytest.mark.parametrize("N", [1, 2, 3, 5])
def test_sigmadelta_matches_numpy_restatement(N):
    rng = np.random.default_rng(N)
    frames = rng.integers(0, 256, size=(9, 37 * 53), dtype=np.uint8)
    # smooth drifting frames too, so that var climbs above 1
    base = rng.integers(0, 200, size=37 * 53)
    drift = np.stack([np.clip(base + 7 * k + rng.integers(-8, 9, base.size), 0, 255) for k in range(9)])
    for fr in (frames, drift.astype(np.uint8)):
        got = O.sigmadelta(fr, N)
        want = sigmadelta_numpy(fr, N)
        for g, w in zip(got[:3], want[:3]):
            assert np.array_equal(g, w)
        assert np.array_equal(got[3].astype(np.int64), want[3])


def test_sigmadelta_demo_sequence():
    """The demo's own input: a 255 rectangle moving by k%10 on a zero
    background (demo/sigmadelta/nt2/sigmadelta_nt2.cpp:79-91), N = 3."""
    h, w, nf = 48, 64, 12
    frames = np.zeros((nf, h, w), np.uint8)
    for k in range(nf):
        i = np.arange(1, h + 1)[:, None]
        j = np.arange(1, w + 1)[None, :]
        m = (i - 1 > h // 4) & (i - 1 < h // 2) & (j - 1 > w // 4 + k % 10) & (j - 1 < w // 2 + k % 10)
        frames[k][m] = 255
    fr = frames.reshape(nf, -1)
    got = O.sigmadelta(fr, 3)
    want = sigmadelta_numpy(fr, 3)
    for g, wv in zip(got[:3], want[:3]):
        assert np.array_equal(g, wv)


def test_sigmadelta_closed_forms():
    # a static scene never moves: var stays 1, labels stay 0
    fr = np.tile(np.arange(256, dtype=np.uint8), (6, 1))
    bkg, var, label, counts = O.sigmadelta(fr, 2)
    assert np.array_equal(bkg, fr[0]) and np.all(var == 1) and np.all(label == 0)
    assert np.all(counts == 0)
    # one pixel stepping from 0 to 255: bkg climbs by one per frame, diff = 255 - k,
    # var climbs by one per frame (it stays below mul = min(255, 2 diff)), label = diff >= var
    nf = 40
    fr = np.zeros((nf, 1), np.uint8)
    fr[1:] = 255
    for k in range(1, nf):
        bkg, var, label, _ = O.sigmadelta(fr[:k + 1], 2)
        assert bkg[0] == k and var[0] == 1 + k and label[0] == int(255 - k >= 1 + k)


def test_fold_kats():
    # globalsum(reshape(1..9, 3, 3)) == 45: modules/core/reduction/unit/table/globalsum.cpp:22-29
    assert O.sum_f32(np.arange(1, 10, dtype=np.float32)) == 45.0
    # sum of ones == M: modules/arch/openmp/unit/functions/sum.cpp:22-56 (here also ragged sizes)
    for m in (1, 7, 8, 9, 63, 64, 65, 1000, 4097):
        assert O.sum_f32(np.ones(m, np.float32)) == float(m)
    # dot(42, 42) == 42^2: modules/core/reduction/unit/table/dot.cpp:22-33
    assert O.dot_f32(np.array([42], np.float32), np.array([42], np.float32)) == 1764.0
    assert O.sum_f32(np.zeros(0, np.float32)) == 0.0


@pytest.mark.parametrize("n", [5, 8, 1000, 1 << 16])
def test_sum_and_dot_error_bounds(n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    y = rng.uniform(-1, 1, n).astype(np.float32)
    eps = np.finfo(np.float32).eps
    s = float(O.sum_f32(x))
    ex = math.fsum(x.astype(np.float64))
    assert abs(s - ex) <= n * eps * float(np.abs(x).sum())
    p = (x * y).astype(np.float32)  # the f32 products NT2 accumulates
    d = float(O.dot_f32(x, y))
    ex = math.fsum(p.astype(np.float64))
    assert abs(d - ex) <= n * eps * float(np.abs(p).sum())
    # exact (double-double) entry points agree with fsum
    assert O.sum_f32_exact(x) == pytest.approx(math.fsum(x.astype(np.float64)), rel=0, abs=1e-12)
    # integer-valued data: every order gives the exact sum
    xi = rng.integers(-1000, 1000, n).astype(np.float32)
    assert float(O.sum_f32(xi)) == float(xi.astype(np.int64).sum())


def test_saxpy_two_roundings():
    # y = a*x + y with a product and a sum rounded separately (no contraction)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, 4099).astype(np.float32)
    y = rng.uniform(-1, 1, 4099).astype(np.float32)
    got = O.saxpy(2.5, x, y)
    want = (np.float32(2.5) * x).astype(np.float32) + y
    assert np.array_equal(got.view(np.uint32), want.astype(np.float32).view(np.uint32))


def _golden():
    d = np.load(os.path.join(HERE, "golden", "bs_golden.npz"))
    return d["S"], d["X"], d["T"], float(d["r"]), float(d["v"]), d["call"], d["put"]


def test_bs_fp64_reference_matches_mpmath():
    S, X, T, r, v, call, put = _golden()
    c, p = O.bs_fp64ref(S, X, T, r, v)
    scale = S.astype(np.float64) + X
    assert np.max(np.abs(c - call) / scale) < 1e-14
    assert np.max(np.abs(p - put) / scale) < 1e-14
    # deep out-of-the-money prices keep their relative accuracy where they are not tiny
    big = np.abs(call) > 1e-6 * scale
    assert np.max(np.abs(c[big] - call[big]) / np.abs(call[big])) < 1e-9


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bs_nt2_evaluation_within_rounding_scale(prec):
    """NT2's textbook evaluation in f32 / f64: absolute error within a few
    rounding units of the S and X*exp(-rT) terms (the cancellation in the
    difference is inherent to the formula)."""
    S, X, T, r, v, call, put = _golden()
    if prec == "f32":
        c, p = O.bs_f32(S, X, T, np.float32(r), np.float32(v))
        c, p = c.astype(np.float64), p.astype(np.float64)
        # golden values at the f32-rounded r and v
        gc, gp = O.bs_fp64ref(S, X, T, float(np.float32(r)), float(np.float32(v)))
        eps = np.finfo(np.float32).eps
    else:
        c, p = O.bs_f64(S.astype(np.float64), X.astype(np.float64), T.astype(np.float64), r, v)
        gc, gp = call, put
        eps = np.finfo(np.float64).eps
    scale = S.astype(np.float64) + X
    assert np.max(np.abs(c - gc) / scale) < 64 * eps
    assert np.max(np.abs(p - gp) / scale) < 64 * eps
    # put-call parity holds to the same scale: C - P = S - X exp(-rT)
    par = S.astype(np.float64) - X.astype(np.float64) * np.exp(-r * T.astype(np.float64))
    assert np.max(np.abs((c - p) - par) / scale) < 128 * eps


def test_bs_demo_mode_formula():
    """The demo's d1 = log(S/X) + ((v^2/2 + r) T)/(v sqrt(T))
    (demo/blackscholes/nt2/blackandscholes_nt2.cpp:56) differs from the
    textbook d1; the demo mode must follow the demo."""
    S, X, T, r, v, call, put = _golden()
    c, _ = O.bs_f32(S, X, T, np.float32(r), np.float32(v), mode=1)
    s, x, t = (a.astype(np.float64) for a in (S, X, T))
    rr, vv = float(np.float32(r)), float(np.float32(v))
    d1 = np.log(s / x) + ((vv * vv * 0.5 + rr) * t) / (vv * np.sqrt(t))
    d2 = d1 - vv * np.sqrt(t)
    from scipy.special import ndtr
    want = s * ndtr(d1) - x * np.exp(-rr * t) * ndtr(d2)
    assert np.max(np.abs(c - want) / (s + x)) < 1e-5  # fastnormcdf is good to ~1e-7 absolute
