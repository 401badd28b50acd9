"""ctypes binding of libnt2h.so (the C ABI declared in include/nt2h.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_NAME = "libnt2h.so"

EXACT, FAST, ACCURATE = 0, 1, 2
BS_TEXTBOOK, BS_DEMO = 0, 1
OPS2 = {"atan2": 0, "pow": 1, "atan2d": 2, "logspace_add": 3, "logspace_sub": 4}
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
H2D, D2H, D2D = 1, 2, 3

_lib = None


class Nt2hError(RuntimeError):
    pass


def lib_path():
    return os.path.join(HERE, LIB_NAME)


def lib():
    """Load libnt2h.so; fail loudly when it is missing (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise Nt2hError("libnt2h.so is not built (expected at %s); run "
                        "`python -c 'import __graft_entry__ as g; g.build()'`" % path)
    L = ctypes.CDLL(path)
    P, S, I, F, D, U64 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                          ctypes.c_float, ctypes.c_double, ctypes.c_uint64)
    sig = {
        "nt2h_abi_version": ([], I),
        "nt2h_last_error": ([], ctypes.c_char_p),
        "nt2h_device_info": ([P, P, P, S], I),
        "nt2h_malloc": ([P, S], I),
        "nt2h_free": ([P], I),
        "nt2h_memcpy": ([P, P, S, I, P], I),
        "nt2h_stream_create": ([P], I),
        "nt2h_stream_destroy": ([P], I),
        "nt2h_stream_sync": ([P], I),
        "nt2h_fill_uniform_f32": ([P, S, D, D, U64, P], I),
        "nt2h_fill_uniform_f64": ([P, S, D, D, U64, P], I),
        "nt2h_fill_normal_f32": ([P, S, U64, P], I),
        "nt2h_fill_frames_u8": ([P, I, I, I, I, I, U64, P], I),
        "nt2h_map_f32": ([I, S, P, P, I, P], I),
        "nt2h_map_f64": ([I, S, P, P, I, P], I),
        "nt2h_sincos_f32": ([S, P, P, P, I, P], I),
        "nt2h_sincos_f64": ([S, P, P, P, I, P], I),
        "nt2h_saxpy_f32": ([S, F, P, P, P], I),
        "nt2h_stream_copy": ([P, P, S, I, P], I),
        "nt2h_sum_f32": ([S, P, P, P], I),
        "nt2h_dot_f32": ([S, P, P, P, P], I),
        "nt2h_dot_f32_noalloc": ([S, P, P, P, P], I),
        "nt2h_blackscholes_f32": ([S, P, P, P, F, F, P, P, I, I, P], I),
        "nt2h_blackscholes_f64": ([S, P, P, P, D, D, P, P, I, P], I),
        "nt2h_bs_set_grid": ([I], I),
        "nt2h_sigmadelta_set_config": ([I], I),
        "nt2h_sigmadelta_u8": ([S, I, P, S, P, P, P, I, P, P], I),
        "nt2h_sigmadelta_step_u8": ([S, P, P, P, P, I, P], I),
        "nt2h_expr4_f32": ([S, P, P, P, I, P], I),
        "nt2h_expr4_sum_f32": ([S, P, P, P, I, P], I),
        "nt2h_sum_dim_f32": ([P, P, I, P, I, P], I),
        "nt2h_sum_dim_f64": ([P, P, I, P, I, P], I),
        "nt2h_cumsum_dim_f32": ([P, P, I, P, I, P], I),
        "nt2h_cumsum_dim_f64": ([P, P, I, P, I, P], I),
        "nt2h_cumprod_dim_f32": ([P, P, I, P, I, P], I),
        "nt2h_cumprod_dim_f64": ([P, P, I, P, I, P], I),
        "nt2h_cummin_dim_f32": ([P, P, I, P, I, P], I),
        "nt2h_cummin_dim_f64": ([P, P, I, P, I, P], I),
        "nt2h_cummax_dim_f32": ([P, P, I, P, I, P], I),
        "nt2h_cummax_dim_f64": ([P, P, I, P, I, P], I),
        "nt2h_rgb2grey_f32": ([S, P, P, P, P, I, P], I),
        "nt2h_rgb2yuv_f32": ([S, P, P, P, P, P, P, I, P], I),
        "nt2h_rgb2grey_u8": ([S, P, P, P, P, P], I),
        "nt2h_map2_f32": ([I, S, P, P, P, I, P], I),
        "nt2h_map2_f64": ([I, S, P, P, P, I, P], I),
        "nt2h_linspace_f32": ([S, F, F, P, I, P], I),
        "nt2h_linspace_f64": ([S, D, D, P, I, P], I),
        "nt2h_mandelbrot_f32": ([S, P, P, P, I, I, P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    _extend(L)
    return L


def _extend(L):
    """Signatures of the batch and server entry points (optional symbols)."""
    P, S, I, D = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
    extra = {
        "nt2h_batch_create": ([P, I, P, P, P, I], I),  # noqa: E501
        "nt2h_batch_run": ([P, P, P], I),
        "nt2h_batch_run_plain": ([P, P, P], I),
        "nt2h_batch_set_config": ([P, I, I, I], I),
        "nt2h_batch_autotune": ([P, P, P], I),
        "nt2h_batch_get_config": ([P, P, P, P], I),
        "nt2h_batch_destroy": ([P], I),
    }
    for name, (args, res) in extra.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res


def check(rc, what=""):
    if rc != 0:
        msg = lib().nt2h_last_error()
        raise Nt2hError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))
    return rc


def ptr(t):
    """Device/host address of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return t.ctypes.data_as(ctypes.c_void_p)


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
