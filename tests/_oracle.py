"""ctypes binding of the oracle (oracle/liboracle.so) for the tests and for
bench.py's CPU-baseline leg.  The oracle is the checker: nothing in the
product path imports this module."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

OPS = {"exp": 0, "log": 1, "sqrt": 2, "erfc": 3, "normcdf": 4,
       "fastnormcdf": 5, "cos": 6, "sin": 7, "abs": 8, "exp2": 9, "exp10": 10, "log2": 11,
       "log10": 12, "floor": 13, "ceil": 14, "trunc": 15, "round": 16, "round2even": 17, "sign": 18,
       "expm1": 19, "log1p": 20, "tanh": 21, "sinh": 22, "cosh": 23, "atan": 24, "tan": 25,
       "asin": 26, "acos": 27, "erf": 28, "cbrt": 29, "asinh": 30, "acosh": 31, "atanh": 32,
       "cot": 33, "sec": 34, "csc": 35, "sinpi": 36, "cospi": 37, "tanpi": 38, "cotpi": 39, "sind": 40,
       "cosd": 41, "tand": 42, "cotd": 43, "coth": 44, "sech": 45, "csch": 46, "acoth": 47, "asech": 48, "acsch": 49,
       "acot": 50, "asec": 51, "acsc": 52, "acotd": 53, "asecd": 54, "acscd": 55, "acosd": 56, "acospi": 57,
       "asind": 58, "asinpi": 59, "atand": 60, "atanpi": 61, "indeg": 62, "inrad": 63, "secd": 64,
       "cscd": 65, "secpi": 66, "cscpi": 67, "sinc": 68, "sincpi": 69,
       "sqrt1pm1": 70, "exprecneg": 71, "exprecnegc": 72, "reallog": 73, "realsqrt": 74, "expx2": 75,
       "erfinv": 76, "erfcinv": 77, "sinhc": 78}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB_PATH)
        P, S = ctypes.c_void_p, ctypes.c_size_t
        _lib.oracle_map_f32.argtypes = [ctypes.c_int, P, P, S]
        _lib.oracle_map_f64.argtypes = [ctypes.c_int, P, P, S]
        _lib.oracle_sincos_f32.argtypes = [P, P, P, S]
        _lib.oracle_sincos_f64.argtypes = [P, P, P, S]
        _lib.oracle_saxpy_f32.argtypes = [S, ctypes.c_float, P, P]
        _lib.oracle_sum_f32.argtypes = [P, S]
        _lib.oracle_sum_f32.restype = ctypes.c_float
        _lib.oracle_sum_f64.argtypes = [P, S]
        _lib.oracle_sum_f64.restype = ctypes.c_double
        for f in ("oracle_dot_f32", "oracle_dot_f32_omp"):
            getattr(_lib, f).argtypes = [P, P, S]
            getattr(_lib, f).restype = ctypes.c_float
        _lib.oracle_dot_f32_exact.argtypes = [P, P, S]
        _lib.oracle_dot_f32_exact.restype = ctypes.c_double
        _lib.oracle_sum_f32_exact.argtypes = [P, S]
        _lib.oracle_sum_f32_exact.restype = ctypes.c_double
        _lib.oracle_bs_f32.argtypes = [S, P, P, P, ctypes.c_float, ctypes.c_float, P, P, ctypes.c_int]
        _lib.oracle_bs_f64.argtypes = [S, P, P, P, ctypes.c_double, ctypes.c_double, P, P]
        _lib.oracle_bs_fp64ref_f32in.argtypes = [S, P, P, P, ctypes.c_double, ctypes.c_double, P, P]
        _lib.oracle_bs_fp64ref_f64in.argtypes = [S, P, P, P, ctypes.c_double, ctypes.c_double, P, P]
        _lib.oracle_bs_quad_f64in.argtypes = [S, P, P, P, ctypes.c_double, ctypes.c_double, P, P, P, P]
        _lib.oracle_u8_binop.argtypes = [ctypes.c_int, P, P, P, S]
        _lib.oracle_sigmadelta_u8.argtypes = [S, ctypes.c_int, P, S, P, P, P, ctypes.c_int, P]
        _lib.oracle_expr4_f32.argtypes = [P, P, P, S]
        _lib.oracle_expr4_sum_f32.argtypes = [P, P, S]
        _lib.oracle_expr4_sum_f32.restype = ctypes.c_float
        _lib.oracle_expr4_sum_fp64.argtypes = [P, P, S]
        _lib.oracle_expr4_sum_fp64.restype = ctypes.c_double
        for f in ("oracle_sum_dim_f32", "oracle_sum_dim_f64", "oracle_cumsum_dim_f32", "oracle_cumsum_dim_f64"):
            getattr(_lib, f).argtypes = [P, P, ctypes.c_int, P]
        for f in ("oracle_cumop_dim_f32", "oracle_cumop_dim_f64"):
            getattr(_lib, f).argtypes = [P, P, ctypes.c_int, ctypes.c_int, P]
        _lib.oracle_rgb2grey_f32.argtypes = [S, P, P, P, P]
        _lib.oracle_rgb2yuv_f32.argtypes = [S, P, P, P, P, P, P]
        _lib.oracle_rgb2grey_u8.argtypes = [S, P, P, P, P]
        _lib.oracle_map2_f32.argtypes = [ctypes.c_int, P, P, P, S]
        _lib.oracle_map2_f64.argtypes = [ctypes.c_int, P, P, P, S]
        _lib.oracle_powi_f32.argtypes = [P, P, P, S]
        _lib.oracle_powi_f64.argtypes = [P, P, P, S]
        _lib.oracle_linspace_f32.argtypes = [S, ctypes.c_float, ctypes.c_float, P]
        _lib.oracle_linspace_f64.argtypes = [S, ctypes.c_double, ctypes.c_double, P]
        _lib.oracle_mandelbrot_f32.argtypes = [S, P, P, P, ctypes.c_int]
        _lib.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib.oracle_get_threads.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_ph_fix(on):
    lib().oracle_set_ph_fix(int(bool(on)))


def set_threads(n):
    lib().oracle_set_threads(int(n))


def threads():
    return lib().oracle_get_threads()


def map_f32(op, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    assert lib().oracle_map_f32(OPS[op], _p(x), _p(y), x.size) == 0
    return y


def map_f64(op, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    assert lib().oracle_map_f64(OPS[op], _p(x), _p(y), x.size) == 0
    return y


def sincos(x):
    x = np.ascontiguousarray(x)
    s, c = np.empty_like(x), np.empty_like(x)
    f = lib().oracle_sincos_f32 if x.dtype == np.float32 else lib().oracle_sincos_f64
    assert f(_p(x), _p(s), _p(c), x.size) == 0
    return s, c


def saxpy(a, x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.array(y, dtype=np.float32, copy=True)
    lib().oracle_saxpy_f32(x.size, a, _p(x), _p(y))
    return y


def sum_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.float32(lib().oracle_sum_f32(_p(x), x.size))


def dot_f32(x, y, omp=False):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    f = lib().oracle_dot_f32_omp if omp else lib().oracle_dot_f32
    return np.float32(f(_p(x), _p(y), x.size))


def dot_f32_exact(x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    return lib().oracle_dot_f32_exact(_p(x), _p(y), x.size)


def sum_f32_exact(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return lib().oracle_sum_f32_exact(_p(x), x.size)


def bs_f32(S, X, T, r, v, mode=0):
    S, X, T = (np.ascontiguousarray(a, dtype=np.float32) for a in (S, X, T))
    c, p = np.zeros_like(S), np.zeros_like(S)
    assert lib().oracle_bs_f32(S.size, _p(S), _p(X), _p(T), r, v, _p(c), _p(p), mode) == 0
    return c, p


def bs_f64(S, X, T, r, v):
    S, X, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (S, X, T))
    c, p = np.zeros_like(S), np.zeros_like(S)
    assert lib().oracle_bs_f64(S.size, _p(S), _p(X), _p(T), r, v, _p(c), _p(p)) == 0
    return c, p


def bs_fp64ref(S, X, T, r, v):
    if S.dtype == np.float32:
        S, X, T = (np.ascontiguousarray(a, dtype=np.float32) for a in (S, X, T))
        f = lib().oracle_bs_fp64ref_f32in
    else:
        S, X, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (S, X, T))
        f = lib().oracle_bs_fp64ref_f64in
    c = np.zeros(S.size)
    p = np.zeros(S.size)
    assert f(S.size, _p(S), _p(X), _p(T), r, v, _p(c), _p(p)) == 0
    return c, p


def bs_quad(S, X, T, r, v):
    """Quad-precision Black-Scholes of f64 options: (call_hi, call_lo,
    put_hi, put_lo), each price the double-double hi + lo."""
    S, X, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (S, X, T))
    out = [np.zeros(S.size) for _ in range(4)]
    assert lib().oracle_bs_quad_f64in(S.size, _p(S), _p(X), _p(T), r, v, *(_p(o) for o in out)) == 0
    return tuple(out)


def u8_binop(op, a, b):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    y = np.empty_like(a)
    assert lib().oracle_u8_binop(op, _p(a), _p(b), _p(y), a.size) == 0
    return y


def sigmadelta(frames, N, bkg=None, var=None, label=None):
    """frames: uint8 array [nframes, npix]; the state starts from bkg, var and
    label (by default the demo's: frame 0, ones, zeros).  Returns bkg, var,
    label, counts."""
    frames = np.ascontiguousarray(frames, dtype=np.uint8)
    nf, npix = frames.shape
    bkg = frames[0].copy() if bkg is None else np.array(bkg, dtype=np.uint8)
    var = np.ones(npix, dtype=np.uint8) if var is None else np.array(var, dtype=np.uint8)
    label = np.zeros(npix, dtype=np.uint8) if label is None else np.array(label, dtype=np.uint8)
    counts = np.zeros(nf, dtype=np.uint32)
    assert lib().oracle_sigmadelta_u8(npix, nf, _p(frames), npix, _p(bkg), _p(var),
                                      _p(label), N, _p(counts)) == 0
    return bkg, var, label, counts


def expr4(x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    z = np.empty_like(x)
    assert lib().oracle_expr4_f32(_p(x), _p(y), _p(z), x.size) == 0
    return z


def expr4_sum(x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    return np.float32(lib().oracle_expr4_sum_f32(_p(x), _p(y), x.size))


def expr4_sum_fp64(x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    return lib().oracle_expr4_sum_fp64(_p(x), _p(y), x.size)


def _ext4(ext):
    e = list(ext) + [1] * (4 - len(ext))
    return np.array(e, dtype=np.uint64)


def sum_dim(x, ext, dim):
    """NT2's sum(a, dim) of the flat column-major table x of extent ext."""
    x = np.ascontiguousarray(x)
    e = _ext4(ext)
    oe = e.copy()
    if dim <= 4:
        oe[dim - 1] = 1
    out = np.empty(int(np.prod(oe)), dtype=x.dtype)
    f = lib().oracle_sum_dim_f32 if x.dtype == np.float32 else lib().oracle_sum_dim_f64
    assert f(_p(x), _p(e), dim, _p(out)) == 0
    return out


def cumsum_dim(x, ext, dim):
    """NT2's cumsum(a, dim) of the flat column-major table x of extent ext."""
    x = np.ascontiguousarray(x)
    e = _ext4(ext)
    out = np.empty_like(x)
    f = lib().oracle_cumsum_dim_f32 if x.dtype == np.float32 else lib().oracle_cumsum_dim_f64
    assert f(_p(x), _p(e), dim, _p(out)) == 0
    return out


CUMOPS = {"cumprod": 1, "cummin": 2, "cummax": 3}


def cumop_dim(x, ext, dim, op):
    """NT2's cumprod / cummin / cummax(a, dim) of the flat column-major table x."""
    if op == "cumsum":
        return cumsum_dim(x, ext, dim)
    x = np.ascontiguousarray(x)
    e = _ext4(ext)
    out = np.empty_like(x)
    f = lib().oracle_cumop_dim_f32 if x.dtype == np.float32 else lib().oracle_cumop_dim_f64
    assert f(_p(x), _p(e), dim, CUMOPS[op], _p(out)) == 0
    return out


def rgb2grey(r, g, b):
    """gr = 0.3 r + 0.59 g + 0.11 b (float32, each operation rounded)."""
    r, g, b = (np.ascontiguousarray(v, dtype=np.float32) for v in (r, g, b))
    out = np.empty_like(r)
    assert lib().oracle_rgb2grey_f32(r.size, _p(r), _p(g), _p(b), _p(out)) == 0
    return out


def rgb2grey_u8(r, g, b):
    """min(255, (77 r + 151 g + 28 b + 128) >> 8) over uint8 planes."""
    r, g, b = (np.ascontiguousarray(v, dtype=np.uint8) for v in (r, g, b))
    out = np.empty_like(r)
    assert lib().oracle_rgb2grey_u8(r.size, _p(r), _p(g), _p(b), _p(out)) == 0
    return out


def rgb2yuv(r, g, b):
    r, g, b = (np.ascontiguousarray(v, dtype=np.float32) for v in (r, g, b))
    y, u, v = np.empty_like(r), np.empty_like(r), np.empty_like(r)
    assert lib().oracle_rgb2yuv_f32(r.size, _p(r), _p(g), _p(b), _p(y), _p(u), _p(v)) == 0
    return y, u, v


def linspace(lo, hi, n, dtype=np.float32):
    out = np.empty(n, dtype=dtype)
    f = lib().oracle_linspace_f32 if dtype == np.float32 else lib().oracle_linspace_f64
    assert f(n, lo, hi, _p(out)) == 0
    return out


def mandelbrot(a, b, max_iter=256):
    """C = arrayfun(step, A, B) over flat float32 tables (NT2's packs of 8)."""
    a, b = (np.ascontiguousarray(v, dtype=np.float32) for v in (a, b))
    c = np.empty(a.size, dtype=np.int32)
    assert lib().oracle_mandelbrot_f32(a.size, _p(a), _p(b), _p(c), int(max_iter)) == 0
    return c


def atan2(y, x):
    """NT2's atan2(y, x) over flat tables (packs of 8 / 4, scalar tail)."""
    y = np.ascontiguousarray(y)
    x = np.ascontiguousarray(x, dtype=y.dtype)
    out = np.empty_like(y)
    f = lib().oracle_map2_f32 if y.dtype == np.float32 else lib().oracle_map2_f64
    assert f(0, _p(y), _p(x), _p(out), y.size) == 0
    return out


def pow_int(a, k):
    """NT2's pow with an integral exponent over flat tables (repeated squaring)."""
    a = np.ascontiguousarray(a)
    k = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.int64), a.shape))
    out = np.empty_like(a)
    f = lib().oracle_powi_f32 if a.dtype == np.float32 else lib().oracle_powi_f64
    assert f(_p(a), _p(k), _p(out), a.size) == 0
    return out


def pow(a, b):
    """NT2's pow(a, b) over flat tables (packs of 8 / 4, scalar tail)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty_like(a)
    f = lib().oracle_map2_f32 if a.dtype == np.float32 else lib().oracle_map2_f64
    assert f(1, _p(a), _p(b), _p(out), a.size) == 0
    return out


def atan2d(y, x):
    """NT2's atan2d(y, x) = indeg(atan2(y, x)) over flat tables."""
    y = np.ascontiguousarray(y)
    x = np.ascontiguousarray(x, dtype=y.dtype)
    out = np.empty_like(y)
    f = lib().oracle_map2_f32 if y.dtype == np.float32 else lib().oracle_map2_f64
    assert f(2, _p(y), _p(x), _p(out), y.size) == 0
    return out


def logspace(op, a, b):
    """NT2's logspace_add / logspace_sub(a, b) over flat tables."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty_like(a)
    f = lib().oracle_map2_f32 if a.dtype == np.float32 else lib().oracle_map2_f64
    assert f({"logspace_add": 3, "logspace_sub": 4}[op], _p(a), _p(b), _p(out), a.size) == 0
    return out
