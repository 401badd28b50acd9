"""Pinning the oracle's sum(a, dim), cumsum(a, dim), cumprod, cummin and
cummax (CPU only).

* NT2's own unit-test known answers: modules/core/reduction/unit/table/sum.cpp
  (sum_1D, sum_2D) and modules/core/swar/unit/table/cumsum.cpp (cumsum_inner,
  cumsum_outer on reshape(1..30, 3, 5, 2));
* the order NT2's SIMD inner fold fixes, on a column where every other order
  gives a different float;
* invariants: exact-arithmetic error bounds against math.fsum, the last
  element of a sequential scan equals the sequential fold, sums over singleton
  dimensions copy the table.
"""
import math

import numpy as np
import pytest

import _oracle as O

KAT = np.arange(1, 31, dtype=np.float64)  # reshape(_(1, 30), 3, 5, 2), column-major


def test_nt2_cumsum_kats():
    # modules/core/swar/unit/table/cumsum.cpp:20-45 (cumsum(a)) and :47-72 (cumsum(a, 2))
    inner = [1, 3, 6, 4, 9, 15, 7, 15, 24, 10, 21, 33, 13, 27, 42,
             16, 33, 51, 19, 39, 60, 22, 45, 69, 25, 51, 78, 28, 57, 87]
    outer = [1, 2, 3, 5, 7, 9, 12, 15, 18, 22, 26, 30, 35, 40, 45,
             16, 17, 18, 35, 37, 39, 57, 60, 63, 82, 86, 90, 110, 115, 120]
    assert list(O.cumsum_dim(KAT, (3, 5, 2), 1)) == inner
    assert list(O.cumsum_dim(KAT, (3, 5, 2), 2)) == outer
    assert list(O.cumsum_dim(KAT.astype(np.float32), (3, 5, 2), 1)) == inner


def test_nt2_sum_kats():
    # modules/core/reduction/unit/table/sum.cpp: sum_1D and sum_2D
    M, N = 5, 4
    a = np.ones(M)
    assert list(O.sum_dim(a, (M,), 1)) == [M]
    for j in (2, 3, 4, 5):  # sum over a singleton dimension returns a
        assert list(O.sum_dim(a, (M,), j)) == list(a)
    assert list(O.sum_dim(2 * a, (M,), 1)) == [2 * M]  # sum(a + b)
    a2 = np.ones(M * N)
    assert list(O.sum_dim(a2, (M, N), 1)) == [M] * N
    assert list(O.sum_dim(a2, (M, N), 2)) == [N] * M
    for j in (3, 4):
        assert list(O.sum_dim(a2, (M, N), j)) == list(a2)
    # sum of an empty table is 0: r = sum(c); r(1) == 0
    assert list(O.sum_dim(np.zeros(0), (0,), 1)) == [0.0]


def test_inner_fold_pack_order():
    """One pack of 8: NT2 sums the lanes as ((l1+l5)+(l3+l7)) + ((l2+l6)+(l0+l4))
    (avx/all_reduce.hpp:52-66).  With l0 = 2^24 and ones elsewhere that is
    4 + (2 + 2^24) = 2^24 + 6, where the sequential order gives 2^24 and a
    pairwise one 2^24 + 4 or 2^24 + 8."""
    col = np.array([2.0 ** 24, 1, 1, 1, 1, 1, 1, 1], dtype=np.float32)
    assert O.sum_dim(col, (8,), 1)[0] == np.float32(2.0 ** 24 + 6)
    # a row vector (1 x 8) folds along dim 2 with the same inner fold (its
    # first non-singleton dimension), a 8 x 1 column along dim 2 is a copy
    assert O.sum_dim(col, (1, 8), 2)[0] == np.float32(2.0 ** 24 + 6)
    assert np.array_equal(O.sum_dim(col, (8, 1), 2), col)
    # the scalar tail is added after the horizontal sum: 9th element
    col9 = np.append(col, np.float32(1))
    assert O.sum_dim(col9, (9,), 1)[0] == np.float32(2.0 ** 24 + 6) + np.float32(1)
    # across packs each lane accumulates its own column of the packs
    two = np.concatenate([col, np.full(8, np.float32(1))])
    # lanes: l0 = 2^24 + 1 -> 2^24 (tie to even), the others 2;
    # hsum: ((2+2)+(2+2)) + ((2+2)+(2^24+2)) = 8 + (2^24 + 6) = 2^24 + 14
    assert O.sum_dim(two, (16,), 1)[0] == np.float32(2.0 ** 24 + 14)


SHAPES = [(37,), (5, 7), (16, 3), (3, 16), (1, 33), (4, 5, 6), (2, 3, 4, 5), (17, 1, 9), (0, 4), (4, 0)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sum_and_cumsum_bounds(shape, dtype):
    rng = np.random.default_rng(len(shape) * 100 + sum(shape))
    n = int(np.prod(shape))
    x = rng.uniform(-1, 1, n).astype(dtype)
    a = x.reshape(shape, order="F")
    eps = np.finfo(dtype).eps
    for dim in range(1, 6):
        s = O.sum_dim(x, shape, dim)
        c = O.cumsum_dim(x, shape, dim)
        ax = dim - 1
        if ax < len(shape):
            ref = np.apply_along_axis(lambda v: math.fsum(v.astype(np.float64)), ax, a) if n else \
                np.zeros([1 if k == ax else m for k, m in enumerate(shape)])
            mag = np.sum(np.abs(a.astype(np.float64)), axis=ax)
            L = shape[ax]
            got = s.reshape(ref.shape if n else ref.shape, order="F")
            assert np.all(np.abs(got - ref) <= L * eps * mag + 1e-300)
            cref = np.cumsum(a.astype(np.float64), axis=ax)
            cmag = np.cumsum(np.abs(a.astype(np.float64)), axis=ax)
            got_c = c.reshape(shape, order="F")
            assert np.all(np.abs(got_c - cref) <= L * eps * cmag + 1e-300)
            if n:
                # the scan's last element along the dimension is its sequential sum
                last = np.take(got_c, [L - 1], axis=ax) if L else None
                if L and dim > 1 and int(np.prod(shape[:ax])) != 1:
                    assert np.array_equal(np.squeeze(last, axis=ax), got)  # outer fold is sequential too
        else:
            # a dimension beyond the extent has size 1: both are copies (0 + x)
            assert np.array_equal(s, x + dtype(0))
            assert np.array_equal(c, x + dtype(0))


def test_nt2_cumprod_kats():
    # modules/core/swar/unit/table/cumprod.cpp:20-44: ones(3, 5, 2) is its own
    # cumprod along 1 and 2; cumprod of the row 1..5 along 2 is 1 2 6 24 120,
    # along 1 the row itself
    a = np.ones(30)
    assert list(O.cumop_dim(a, (3, 5, 2), 1, "cumprod")) == list(a)
    assert list(O.cumop_dim(a, (3, 5, 2), 2, "cumprod")) == list(a)
    aa = np.arange(1, 6, dtype=np.float64)
    assert list(O.cumop_dim(aa, (1, 5), 2, "cumprod")) == [1, 2, 6, 24, 120]
    assert list(O.cumop_dim(aa, (1, 5), 1, "cumprod")) == list(aa)
    assert list(O.cumop_dim(aa.astype(np.float32), (1, 5), 2, "cumprod")) == [1, 2, 6, 24, 120]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,dim", [((1000,), 1), ((7, 300), 1), ((7, 300), 2), ((4, 5, 6), 2), ((4, 5, 6), 3)])
def test_cumulative_against_numpy_accumulate(shape, dim, dtype):
    """NaN-free tables without zeros: NT2's sequential product equals numpy's
    multiply.accumulate in the same precision (one rounding per step, in
    order), and its min / max chains equal minimum / maximum.accumulate."""
    rng = np.random.default_rng(len(shape) * 10 + dim)
    n = int(np.prod(shape))
    x = (rng.uniform(0.5, 1.5, n) * rng.choice([-1, 1], n)).astype(dtype)
    a = x.reshape(shape, order="F")
    ax = dim - 1
    for op, ref in (("cumprod", np.multiply.accumulate), ("cummin", np.minimum.accumulate),
                    ("cummax", np.maximum.accumulate)):
        want = ref(a, axis=ax, dtype=dtype).ravel(order="F")
        got = O.cumop_dim(x, shape, dim, op)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (op, shape, dim)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cummin_cummax_nan_and_signed_zero(dtype):
    """NT2's scalar min / max (s < x ? s : x, s > x ? s : x): a NaN element is
    kept at its own position and the chain restarts from the next element;
    between +0 and -0 the later element wins.  fmin/fmax or a NaN-propagating
    min would give other values at every position marked here."""
    nan = np.nan
    x = np.array([3.0, nan, 5.0, 1.0, 2.0, nan, nan, -4.0], dtype)
    mn = O.cumop_dim(x, (8,), 1, "cummin")
    mx = O.cumop_dim(x, (8,), 1, "cummax")
    assert np.array_equal(mn, np.array([3, nan, 5, 1, 1, nan, nan, -4], dtype), equal_nan=True)
    assert np.array_equal(mx, np.array([3, nan, 5, 5, 5, nan, nan, -4], dtype), equal_nan=True)
    z = np.array([0.0, -0.0, 0.0, -0.0], dtype)
    assert np.signbit(O.cumop_dim(z, (4,), 1, "cummin")).tolist() == [False, True, False, True]
    assert np.signbit(O.cumop_dim(z, (4,), 1, "cummax")).tolist() == [False, True, False, True]
    inf = np.array([np.inf, -np.inf, 1.0], dtype)
    assert O.cumop_dim(inf, (3,), 1, "cummin").tolist() == [np.inf, -np.inf, -np.inf]
    assert O.cumop_dim(inf, (3,), 1, "cummax").tolist() == [np.inf, np.inf, np.inf]
