"""sum(a, dim) and cumsum(a, dim) on the GPU through the C ABI
(nt2h_sum_dim_*, nt2h_cumsum_dim_*) against the oracle.

Exact policy: bit-identical to the oracle's restatement of NT2's order for
every shape and dimension (inner fold with pack accumulators, sequential outer
folds and scans).  Fast policy: within the error of one final rounding plus
FP64 accumulation of the exact sum (math.fsum), deterministic run to run.
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import _oracle as O  # noqa: E402
from _gpu import same_bits, say  # noqa: E402
from nt2_amd import EXACT, FAST, ops  # noqa: E402

SHAPES = [(37,), (1000,), (5, 7), (16, 3), (3, 16), (1, 33), (33, 1), (64, 100), (4, 5, 6), (2, 3, 4, 5),
          (17, 1, 9), (0, 4), (4, 0), (9, 2, 1, 3)]


def _dev(x):
    return torch.from_numpy(x).cuda()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", SHAPES)
def test_exact_matches_oracle(shape, dtype):
    rng = np.random.default_rng(sum(shape) + 7 * len(shape))
    n = int(np.prod(shape))
    x = rng.uniform(-1, 1, n).astype(dtype) * np.where(rng.random(n) < 0.1, 1e6, 1).astype(dtype)
    xd = _dev(x)
    for dim in range(1, 6):
        say("dimwise exact", shape, dtype.__name__, "dim=%d" % dim)
        s = ops.sum_dim(xd, shape, dim, EXACT).cpu().numpy()
        assert same_bits(s, O.sum_dim(x, shape, dim)), (shape, dim)
        c = ops.cumsum_dim(xd, shape, dim, EXACT).cpu().numpy()
        assert same_bits(c, O.cumsum_dim(x, shape, dim)), (shape, dim)


def test_exact_nt2_kats():
    # modules/core/swar/unit/table/cumsum.cpp and modules/core/reduction/unit/table/sum.cpp
    kat = np.arange(1, 31, dtype=np.float64)
    c1 = ops.cumsum_dim(_dev(kat), (3, 5, 2), 1, EXACT).cpu().numpy()
    c2 = ops.cumsum_dim(_dev(kat), (3, 5, 2), 2, EXACT).cpu().numpy()
    assert list(c1[:6]) == [1, 3, 6, 4, 9, 15] and c1[-1] == 87
    assert list(c2[:6]) == [1, 2, 3, 5, 7, 9] and c2[-1] == 120
    ones = _dev(np.ones(20, dtype=np.float64))
    assert list(ops.sum_dim(ones, (5, 4), 1, EXACT).cpu().numpy()) == [5] * 4
    assert list(ops.sum_dim(ones, (5, 4), 2, EXACT).cpu().numpy()) == [4] * 5
    col = np.array([2.0 ** 24, 1, 1, 1, 1, 1, 1, 1], dtype=np.float32)
    assert ops.sum_dim(_dev(col), (8,), 1, EXACT).item() == 2.0 ** 24 + 6


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,dim", [((1 << 20,), 1), ((3, 100000), 2), ((20000, 7), 1), ((1, 300000), 2),
                                       ((50000, 3), 1), ((8, 8, 4096), 3), ((3,), 1), ((100, 5000), 1),
                                       ((300, 70000), 1), ((70000, 300), 2)])
def test_fast_within_bounds(shape, dim, dtype):
    rng = np.random.default_rng(len(shape) + dim)
    n = int(np.prod(shape))
    x = rng.uniform(-1, 1, n).astype(dtype)
    say("dimwise fast", shape, dtype.__name__, "dim=%d" % dim)
    xd = _dev(x)
    s1 = ops.sum_dim(xd, shape, dim, FAST)
    s2 = ops.sum_dim(xd, shape, dim, FAST)
    assert torch.equal(s1, s2)  # deterministic
    a = x.reshape(shape, order="F").astype(np.float64)
    ax = dim - 1
    ref = np.apply_along_axis(lambda v: math.fsum(v), ax, a).ravel(order="F")
    got = s1.cpu().numpy().astype(np.float64)
    eps = np.finfo(dtype).eps
    mag = np.sum(np.abs(a), axis=ax).ravel(order="F")
    # one rounding of the result plus the accumulation error of compensated sums
    assert np.all(np.abs(got - ref) <= eps * np.abs(ref) + 1e-12 * mag + 1e-300)
    c = ops.cumsum_dim(xd, shape, dim, FAST)
    cref = np.cumsum(a, axis=ax).ravel(order="F")
    cmag = np.cumsum(np.abs(a), axis=ax).ravel(order="F")
    err = np.abs(c.cpu().numpy().astype(np.float64) - cref)
    tol = (eps * np.abs(cref) + 1e-12 * cmag) if dtype == np.float32 else (eps * np.abs(cref) + shape[ax] * eps * cmag)
    assert np.all(err <= tol + 1e-300)


def test_cumsum_in_place():
    x = np.random.default_rng(3).uniform(-1, 1, 3 * 20000).astype(np.float32)
    for pol, shape, dim in ((EXACT, (3, 20000), 2), (EXACT, (20000, 3), 1), (FAST, (20000, 3), 1),
                            (FAST, (3, 20000), 2), (FAST, (300, 200), 1)):
        xd = _dev(x.copy())
        want = ops.cumsum_dim(_dev(x), shape, dim, pol)
        ops.cumsum_dim(xd, shape, dim, pol, out=xd)
        assert torch.equal(xd, want)


def test_argument_errors():
    from nt2_amd.native import Nt2hError
    xd = _dev(np.ones(8, dtype=np.float32))
    with pytest.raises(Nt2hError):
        ops.sum_dim(xd, (8,), 0, EXACT)
    with pytest.raises(Nt2hError):
        ops.sum_dim(xd, (8,), 1, 7)


def _cum_inputs(op, n, dtype, rng):
    if op == "cumprod":  # magnitudes near 1: long chains stay finite
        x = rng.uniform(0.97, 1.03, n) * rng.choice([-1, 1], n)
    else:
        x = rng.uniform(-1, 1, n) * np.where(rng.random(n) < 0.1, 1e6, 1)
    x = x.astype(dtype)
    k = max(1, n // 50)
    idx = rng.integers(0, max(n, 1), k) if n else np.zeros(0, int)
    x[idx[: k // 3]] = np.nan
    x[idx[k // 3: 2 * k // 3]] = 0.0
    x[idx[2 * k // 3:]] = -0.0
    return x


@pytest.mark.parametrize("op", ["cumprod", "cummin", "cummax"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", SHAPES + [(3, 20000), (20000, 3)])
def test_cumulative_matches_oracle(shape, dtype, op):
    """cumprod / cummin / cummax: NT2's order under both policies, bit for bit
    (NaN, +0 and -0 sprinkled in; (3, 20000) along 1 takes the many-lines
    kernel, (20000, 3) along 1 the workgroup-per-line one, dim 2 the strided)."""
    rng = np.random.default_rng(sum(shape) + 11 * len(shape))
    n = int(np.prod(shape))
    x = _cum_inputs(op, n, dtype, rng)
    xd = _dev(x)
    for dim in range(1, 4):
        want = O.cumop_dim(x, shape, dim, op)
        for pol in (EXACT, FAST):
            say("cumulative", op, shape, dtype.__name__, "dim=%d pol=%d" % (dim, pol))
            got = ops.cumsum_dim(xd, shape, dim, pol, op=op).cpu().numpy()
            assert same_bits(got, want), (op, shape, dim, pol)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("op", ["cumsum", "cumprod", "cummin", "cummax"])
@pytest.mark.parametrize("shape", [(200, 20000), (1000, 2500), (2500, 1000)])
def test_scan_many_long_lines(shape, op, dtype):
    """Along dimension 1: 20000 contiguous lines of 200 (the 64-lines-per-wave
    kernel: 3 full tiles of 64 floats or 6 of 32 doubles, then a tail; 20000
    lines leave a partial wave), 2500 lines of 1000 (the 8-lines-per-wave
    kernel: one full tile of 512 floats or 3 of 256 doubles, then a tail; 2500
    lines leave a partial wave) and 1000 lines of 2500 (a workgroup per line);
    along dimension 2 the strided chains (64 loads in flight per chain).
    NT2's order bit for bit: cumsum under the exact policy, the other scans
    under both."""
    rng = np.random.default_rng(sum(shape) + len(op))
    n = int(np.prod(shape))
    x = _cum_inputs("cumsum" if op == "cumsum" else op, n, dtype, rng)
    if op == "cumsum":
        x = np.nan_to_num(x)
    xd = _dev(x)
    for dim in (1, 2):
        if op == "cumsum":
            want, pols = O.cumsum_dim(x, shape, dim), (EXACT,)
        else:
            want, pols = O.cumop_dim(x, shape, dim, op), (EXACT, FAST)
        for pol in pols:
            say("scan many long lines", op, shape, dtype.__name__, "dim=%d pol=%d" % (dim, pol))
            got = ops.cumsum_dim(xd, shape, dim, pol, op=op).cpu().numpy()
            assert same_bits(got, want), (op, shape, dim, pol)


def test_cumulative_in_place_and_errors():
    from nt2_amd.native import Nt2hError
    x = np.random.default_rng(4).uniform(0.5, 1.5, 3 * 20000).astype(np.float32)
    for op in ("cumprod", "cummin", "cummax"):
        want = ops.cumsum_dim(_dev(x), (3, 20000), 1, EXACT, op=op)
        xd = _dev(x.copy())
        ops.cumsum_dim(xd, (3, 20000), 1, EXACT, out=xd, op=op)
        assert torch.equal(xd, want)
        with pytest.raises(Nt2hError):
            ops.cumsum_dim(_dev(x), (3, 20000), 0, EXACT, op=op)
        with pytest.raises(Nt2hError):
            ops.cumsum_dim(_dev(x), (3, 20000), 1, 7, op=op)
