"""The C++ front end (nt2/table.hpp): NT2 demo statements compiled unchanged
with hipcc, run on the GPU, compared with the oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import _oracle as O  # noqa: E402
from _gpu import first_diff, same_bits, say  # noqa: E402

EXE = os.path.join(os.path.dirname(__file__), "cpp", "bin", "frontend_demos")


@pytest.fixture(scope="module")
def outputs(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("front-end demo program not built (run the build)")
    d = str(tmp_path_factory.mktemp("fe"))
    say("frontend_demos")
    p = subprocess.run([EXE, d, "100003"], capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout + p.stderr

    def rd(name, dt):
        return np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)
    return rd


def test_saxpy_statement(outputs):
    x, y0, y = outputs("x", np.float32), outputs("y", np.float32), outputs("saxpy", np.float32)
    assert same_bits(y, O.saxpy(2.5, x, y0))


def test_sdot_and_dot_statements(outputs):
    x, y = outputs("x", np.float32), outputs("saxpy", np.float32)
    p = (x * y).astype(np.float32)
    ex = O.sum_f32_exact(p)
    mag = O.sum_f32_exact(np.abs(p))
    for name in ("sdot", "dot"):
        r = float(outputs(name, np.float32)[0])
        # result rounded to float once, from a sum accurate to ~2^-40 of the magnitude
        assert abs(r - ex) <= abs(float(np.float32(ex)) - ex) + 2 ** -40 * mag, (name, r, ex)


def test_blackscholes_demo_statements(outputs):
    S, X, T = (outputs(n, np.float32) for n in ("S", "X", "T"))
    R = outputs("bs_demo", np.float32)
    want, _ = O.bs_f32(S, X, T, 0.02, 0.30, mode=1)
    assert same_bits(R, want), first_diff(R, want)


def test_blackscholes_tie_statement(outputs):
    S, X, T = (outputs(n, np.float32) for n in ("S", "X", "T"))
    c, p = outputs("bs_call", np.float32), outputs("bs_put", np.float32)
    wc, wp = O.bs_f32(S, X, T, 0.02, 0.30, mode=0)
    assert same_bits(c, wc), first_diff(c, wc)
    assert same_bits(p, wp), first_diff(p, wp)


def test_trig_expression_and_globalsum(outputs):
    xn, yn, z = outputs("xn", np.float32), outputs("yn", np.float32), outputs("expr4", np.float32)
    want = O.expr4(xn, yn)
    assert same_bits(z, want), first_diff(z, want)
    s = float(outputs("expr4_sum", np.float32)[0])
    ex = O.sum_f32_exact(want)
    assert abs(s - ex) <= abs(float(np.float32(ex)) - ex) + 2 ** -40 * O.sum_f32_exact(np.abs(want))


def test_sigmadelta_statements(outputs):
    frames = outputs("frames", np.uint8).reshape(8, 37 * 29)
    b, v, l, _ = O.sigmadelta(frames, 3)  # the demo multiplies by 3 with adds
    assert np.array_equal(outputs("sd_bkg", np.uint8), b)
    assert np.array_equal(outputs("sd_var", np.uint8), v)
    assert np.array_equal(outputs("sd_label", np.uint8), l)
    # globalsum of an integer table: the sum in the element type, wrapping
    assert outputs("frame1_sum", np.uint8)[0] == int(frames[1].astype(np.uint64).sum()) % 256


def test_dimwise_kats(outputs):
    """sum and cumsum along dimensions of reshape(_(1, 30), 3, 5, 2): NT2's
    unit-test answers (modules/core/swar/unit/table/cumsum.cpp,
    modules/core/reduction/unit/table/sum.cpp)."""
    a = np.arange(1, 31, dtype=np.float64)
    assert list(outputs("cumsum_inner", np.float64)) == list(O.cumsum_dim(a, (3, 5, 2), 1))
    assert list(outputs("cumsum_outer", np.float64)[:6]) == [1, 2, 3, 5, 7, 9]
    assert list(outputs("cumsum_outer", np.float64)) == list(O.cumsum_dim(a, (3, 5, 2), 2))
    assert list(outputs("sum_a", np.float64)) == list(O.sum_dim(a, (3, 5, 2), 1))
    assert list(outputs("sum_a2", np.float64)) == list(O.sum_dim(a, (3, 5, 2), 2))
    assert list(outputs("sum_a3", np.float64)) == list(O.sum_dim(a, (3, 5, 2), 3))
    assert list(outputs("sum_a4", np.float64)) == list(a)  # a singleton dimension: a copy


def test_dimwise_statements_bitwise(outputs):
    """Front-end sum/cumsum (exact policy) equal the oracle's NT2-order
    restatement bit for bit, expressions included: cos evaluated per pack of
    8 along the columns (inner fold) or the rows (outer fold) as NT2 does."""
    M = outputs("M", np.float32)
    ext = (256, 129)
    assert same_bits(outputs("M_sum1", np.float32), O.sum_dim(M, ext, 1))
    assert same_bits(outputs("M_sum2", np.float32), O.sum_dim(M, ext, 2))
    assert same_bits(outputs("M_cumsum1", np.float32), O.cumsum_dim(M, ext, 1))
    assert same_bits(outputs("M_cumsum2", np.float32), O.cumsum_dim(M, ext, 2))
    # 256 rows: the columns' packs and the rows' packs (256 a multiple of 16)
    # coincide with the table's packs from index 0, so map then fold restates it
    c = O.map_f32("cos", M)
    assert same_bits(outputs("M_sumcos1", np.float32), O.sum_dim(c, ext, 1))
    assert same_bits(outputs("M_sumcos2", np.float32), O.sum_dim(c, ext, 2))
    e = (M * np.float32(2.0)).astype(np.float32) + np.float32(1.0)
    assert same_bits(outputs("M_sumexpr", np.float32), O.sum_dim(e.astype(np.float32), ext, 1))
    v = np.full(1000, M[0], dtype=np.float32)
    assert outputs("v_sum", np.float32)[0] == O.sum_dim(v, (1000,), 1)[0]
    p = (M / np.float32(1.0e4)).astype(np.float32) + np.float32(1.0)
    assert same_bits(outputs("M_cumprod1", np.float32), O.cumop_dim(p, ext, 1, "cumprod"))
    assert same_bits(outputs("M_cummin2", np.float32), O.cumop_dim(M, ext, 2, "cummin"))
    cs = np.array([O.map_f32("cos", np.array([v], np.float32))[0] for v in M], np.float32)
    assert same_bits(outputs("M_cummaxcos1", np.float32), O.cumop_dim(cs, ext, 1, "cummax"))


def test_exponential_rounding_statements(outputs):
    """exp2, exp10, log2, log10, floor, ceil, trunc, round, round2even and sign
    inside fused statements: every function is lane-independent, so the oracle
    maps each one over the whole table and combines in float32/float64."""
    f32 = np.float32
    xn, yn = outputs("xn", f32), outputs("yn", f32)
    a = O.map_f32("round", xn * f32(10)) + O.map_f32("log2", np.abs(yn) + f32(1))
    want = a - O.map_f32("exp10", xn / f32(8))
    got = outputs("ew_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    b = O.map_f32("round2even", yn / f32(16)) * O.map_f32("sign", xn)
    want = b + O.map_f32("exp2", O.map_f32("floor", xn) + O.map_f32("ceil", -xn) / f32(4))
    got = outputs("ew2_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = (O.map_f32("tanh", xn) * O.map_f32("expm1", xn / f32(4))) + O.map_f32("log1p", np.abs(yn))
    got = outputs("ew3_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    a, b = outputs("at_a", f32), outputs("at_b", f32)
    want = O.map_f32("atan", a * f32(3)) + O.atan2(a, b)
    got = outputs("atan_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f32("tan", a) * f32(0.5) - O.map_f32("tan", b * f32(100))
    got = outputs("tan_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f32("asin", a / f32(3)) * O.map_f32("acos", b / f32(10))
    got = outputs("asin_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = O.pow(np.abs(a) + f32(0.5), b / f32(4)) + O.pow(a, O.map_f32("round", b))
    got = outputs("pow_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = O.pow_int(a, 3) + O.pow_int(b, -2)
    got = outputs("powi_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    near1 = (b * f32(0.001)).astype(f32) + f32(1)
    want = (O.pow_int(a, 2) + O.pow_int(near1, 3000000000)) - O.pow_int(a, -3)
    got = outputs("powu_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = (O.map_f32("erf", a) + O.map_f32("cbrt", b * f32(50))) - O.map_f32("asinh", b * a * f32(100))
    got = outputs("erf_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f32("acosh", np.abs(b) * f32(1.0e6) + f32(1)) * O.map_f32("atanh", a / f32(3))
    got = outputs("acosh_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f32("cot", a) + O.map_f32("sec", b) * O.map_f32("csc", a * f32(7))
    got = outputs("cot_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = (O.map_f32("sinpi", a) - O.map_f32("cospi", b)) + \
        O.map_f32("tanpi", a / f32(7)) * O.map_f32("cotpi", b / f32(3))
    got = outputs("sinpi_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = (O.map_f32("sind", a * f32(100)) - O.map_f32("cosd", b * f32(50))) + \
        O.map_f32("tand", a * f32(30)) * O.map_f32("cotd", b * f32(40))
    got = outputs("sind_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    m = O.map_f32
    want = (m("coth", a) + m("sech", b) * m("csch", a * f32(3))) - m("acoth", b * f32(4))
    got = outputs("coth_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = (m("asech", np.abs(a) / f32(4)) + m("acsch", b)) - m("acot", a) * m("asec", b * f32(3))
    got = outputs("asech_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = ((m("acsc", a * f32(2)) + m("acotd", b)) - m("asecd", a * f32(5))) + m("acscd", b * f32(4))
    got = outputs("acsc_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = (m("acosd", a / f32(4)) - m("acospi", b / f32(4))) + m("asind", a / f32(3)) * m("asinpi", b / f32(5))
    got = outputs("acosd_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = ((m("atand", a) + m("atanpi", b)) + m("indeg", a)) - m("inrad", b * f32(100))
    got = outputs("atand_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = (((m("secd", a * f32(90)) + m("cscd", b * f32(45))) - m("secpi", a) * m("cscpi", b / f32(3))) +
            m("sinc", a * f32(4))) + m("sincpi", b)
    got = outputs("secd_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    got = outputs("atan2d_f32", f32)
    assert same_bits(got, O.atan2d(a, b)), first_diff(got, O.atan2d(a, b))
    want = (((m("sqrt1pm1", a) + m("exprecneg", b)) - m("exprecnegc", a * f32(2))) +
            m("reallog", np.abs(b)) * m("realsqrt", np.abs(a))) + m("expx2", b / f32(2))
    got = outputs("sqrt1pm1_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = O.logspace("logspace_add", a, b) - O.logspace("logspace_sub", a * f32(2), (a * f32(2) - np.abs(b)) - f32(0.5))
    got = outputs("logspace_f32", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = np.where((a > 0) & (b < 1), a, b) + np.where((a < -1) | (b > 2), f32(1), f32(0))
    got = outputs("logical_f32", f32)
    assert same_bits(got, want.astype(f32)), first_diff(got, want)
    d = outputs("M", f32).astype(np.float64) / 100.0
    want = (O.map_f64("trunc", d) * O.map_f64("exp2", O.map_f64("ceil", d / 50.0)) +
            O.map_f64("sign", d) * O.map_f64("log10", np.abs(d) + 1.0))
    got = outputs("ew_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f64("sinh", d / 20.0) - O.map_f64("cosh", d / 30.0)
    got = outputs("ew2_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f64("tan", d / 7.0) + O.map_f64("atan", d)
    got = outputs("tan_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    q = outputs("dq", np.float64)
    assert q.size == 1001
    want = O.map_f64("asin", q) - O.map_f64("acos", q * 0.9)
    got = outputs("asin_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = O.pow(q + 2.0, q * 300.0)
    got = outputs("pow_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = O.pow_int(q * 10.0, 7)
    got = outputs("powi_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = (O.map_f64("erf", q * 8.0) + O.map_f64("cbrt", q * 1e5)) - O.map_f64("asinh", q * 1e9)
    got = outputs("erf_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f64("acosh", np.abs(q) * 1e17 + 1.0) * O.map_f64("atanh", q * 0.999)
    got = outputs("acosh_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = O.map_f64("cot", q * 30.0) + O.map_f64("sec", q * 9.0) * O.map_f64("csc", q * 7.0)
    got = outputs("cot_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = (O.map_f64("sinpi", q * 5.0) - O.map_f64("cospi", q * 3.0)) + \
        O.map_f64("tanpi", q / 0.7) * O.map_f64("cotpi", q * 1.3)
    got = outputs("sinpi_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = (O.map_f64("sind", q * 500.0) - O.map_f64("cosd", q * 300.0)) + \
        O.map_f64("tand", q * 200.0) * O.map_f64("cotd", q * 170.0)
    got = outputs("sind_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    m = O.map_f64
    want = (m("coth", q * 3.0) - m("asech", np.abs(q))) + m("acot", q) * m("asec", q * 30.0)
    got = outputs("coth_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = ((m("acotd", q * 7.0) + m("acscd", q * 20.0)) - m("asind", q * 0.9)) + m("atanpi", q * 9.0)
    got = outputs("acotd_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)
    want = ((m("secd", q * 900.0) - m("cscpi", q * 6.0)) + m("sinc", q * 40.0) * m("sincpi", q * 5.0)) + \
        O.atan2d(q, q * 0.5 - 0.1)
    got = outputs("secd_f64", np.float64)
    assert same_bits(got, want), first_diff(got, want)


def test_image_demo_statements(outputs):
    """rgb2grey (demo/rgb2grey/nt2/rgb2grey_nt2.cpp:44) on the demo's ramp and
    on random planes, rgb2yuv's three statements (demo/rgb2yuv/nt2/rgb2yuv_nt2.cpp:30-35)."""
    f32 = np.float32
    ramp = np.arange(96 * 131, dtype=f32)
    assert same_bits(outputs("grey_ramp", f32), O.rgb2grey(ramp, ramp, ramp))
    r, g, b = outputs("pr", f32), outputs("pg", f32), outputs("pb", f32)
    assert same_bits(outputs("grey", f32), O.rgb2grey(r, g, b))
    y, u, v = O.rgb2yuv(r, g, b)
    assert same_bits(outputs("yuv_y", f32), y)
    assert same_bits(outputs("yuv_u", f32), u)
    assert same_bits(outputs("yuv_v", f32), v)


def test_repeated_subexpressions(outputs):
    """Statements that repeat subexpressions over one view (shared, decided at
    compile time after a host test) and over different views of one node type
    (every occurrence evaluated) give the oracle's bits either way; normcdf(x)
    and normcdf(-x) share erfc's magnitude."""
    f32 = np.float32
    a, b = outputs("at_a", f32), outputs("at_b", f32)
    ea, eb = O.map_f32("exp", a), O.map_f32("exp", b)
    c2 = O.map_f32("cos", b * f32(2))
    want = ea * ea + np.sqrt(ea) - c2 / c2
    got = outputs("rep_same", f32)
    assert same_bits(got, want), first_diff(got, want)
    want = ea * eb + c2 - O.map_f32("cos", b * f32(3)) + ea
    got = outputs("rep_mixed", f32)
    assert same_bits(got, want), first_diff(got, want)
    na, nb = O.map_f32("normcdf", a), O.map_f32("normcdf", b)
    nma, nmb = O.map_f32("normcdf", -a), O.map_f32("normcdf", -b)
    for name, want in (("rep_tie1", na + nb), ("rep_tie2", nma - nmb), ("rep_tie3", na * f32(2)),
                       ("rep_tie4", nma)):
        got = outputs(name, f32)
        assert same_bits(got, want), (name, first_diff(got, want))
    for name, z in (("rep_sum_mixed", ea * ea + eb), ("rep_sum_same", ea * ea)):
        ex = O.sum_f32_exact(z)
        mag = O.sum_f32_exact(np.abs(z))
        r = float(outputs(name, f32)[0])
        assert abs(r - ex) <= abs(float(np.float32(ex)) - ex) + 2 ** -40 * mag, (name, r, ex)


def test_mandelbrot_demo_statements(outputs):
    """A = expand_to(linspace(...)), B = expand_to(colvect(linspace(...))) and
    C = arrayfun(step, A, B) with a functor deciding per pack of 8
    (demo/mandelbrot/nt2/mandelbrot_nt2.cpp:86-88, :102)."""
    n = 203
    lin = O.linspace(-1.5, 1.5, n)
    A, B = outputs("mandel_a", np.float32), outputs("mandel_b", np.float32)
    assert same_bits(A, np.repeat(lin, n))  # A(i, j) = lin(j)
    assert same_bits(B, np.tile(lin, n))    # B(i, j) = lin(i)
    C = outputs("mandel", np.int32)
    want = O.mandelbrot(A, B, 256)
    assert np.array_equal(C, want), first_diff(C, want)
    wrapped = np.array([int((C.astype(np.int64) * 1000003).sum()) % 2 ** 32], np.uint32).view(np.int32)[0]
    assert (C.astype(np.int64) * 1000003).sum() > 2 ** 31  # the int32 sum wraps
    assert outputs("mandel_sum", np.int32)[0] == wrapped
    assert same_bits(outputs("linspace100", np.float32), O.linspace(0.0, 1.0, 100))


def test_device_storage_coherence(outputs):
    """Tables keep a host and a device copy with validity flags: host writes
    after device use are uploaded (all elements, not a prefix), device results
    are downloaded on host access, copies are independent, in-place
    statements read the current device copy, resizing starts from zeros."""
    m = 100003
    p = outputs("coherence", np.float32)
    assert list(p) == [2.0 * m, -1.0, float(m), 43.0, 1.0, 3.0 * m, 2.0 * m, 1.0]
    cp = outputs("coherence_cp", np.float32)
    want = np.arange(1, m + 1, dtype=np.float32) + np.float32(1)
    want[6] = np.float32(0.0)  # -1 + 1
    want[0] = np.float32(43.0)
    assert np.array_equal(cp, want)


def test_table_settings(outputs):
    """table<T, S> with NT2's settings: C_index_ tables index from 0,
    matlab_index_ (the default) from 1, _1D/_2D reject extents with more
    dimensions (construction and assignment from an expression)."""
    p = outputs("settings", np.float32)
    # c0(i, j) = 10 i + j (0-based); m1 = 2 c0 read 1-based; d1 = cumsum of reshape(1:6, 2, 3)
    assert list(p) == [0.0, 46.0, 23.0, 11.0, 1.0, 1.0, 55.0]
    m1 = outputs("settings_m1", np.float32).reshape(4, 3).T  # column-major 3x4
    i, j = np.meshgrid(np.arange(3), np.arange(4), indexing="ij")
    assert np.array_equal(m1, (2 * (10 * i + j)).astype(np.float32))
