"""Python-level calls of the C ABI on torch device tensors (test and bench
plumbing).  Every function launches on torch's current stream."""
import ctypes

import torch

from .native import (BS_TEXTBOOK, EXACT, OPS, OPS2, check, lib, ptr, stream_handle)


def _s():
    return stream_handle()


def fill_uniform(n, lo, hi, seed, dtype=torch.float32, device="cuda"):
    out = torch.empty(n, dtype=dtype, device=device)
    f = lib().nt2h_fill_uniform_f32 if dtype == torch.float32 else lib().nt2h_fill_uniform_f64
    check(f(ptr(out), n, lo, hi, seed, _s()), "fill_uniform")
    return out


def fill_normal(n, seed, device="cuda"):
    out = torch.empty(n, dtype=torch.float32, device=device)
    check(lib().nt2h_fill_normal_f32(ptr(out), n, seed, _s()), "fill_normal")
    return out


def fill_frames(h, w, nframes, nrect=8, noise=8, seed=3, device="cuda"):
    out = torch.empty((nframes, h * w), dtype=torch.uint8, device=device)
    check(lib().nt2h_fill_frames_u8(ptr(out), h, w, nframes, nrect, noise, seed, _s()), "fill_frames")
    return out


def map(op, x, policy=EXACT, out=None):
    y = torch.empty_like(x) if out is None else out
    f = lib().nt2h_map_f32 if x.dtype == torch.float32 else lib().nt2h_map_f64
    check(f(OPS[op], x.numel(), ptr(x), ptr(y), policy, _s()), "map " + op)
    return y


def sincos(x, policy=EXACT):
    s, c = torch.empty_like(x), torch.empty_like(x)
    f = lib().nt2h_sincos_f32 if x.dtype == torch.float32 else lib().nt2h_sincos_f64
    check(f(x.numel(), ptr(x), ptr(s), ptr(c), policy, _s()), "sincos")
    return s, c


def stream_copy(src, dst, grid=0):
    check(lib().nt2h_stream_copy(ptr(src), ptr(dst), src.numel() * src.element_size(), grid, _s()),
          "stream_copy")


def map2(op, x, y, policy=EXACT, out=None):
    """out = f(x, y) elementwise for a binary function (atan2, pow)."""
    if y.numel() != x.numel() or y.dtype != x.dtype or not (x.is_contiguous() and y.is_contiguous()):
        raise ValueError("map2 %s: x and y must be contiguous tables of one size and type" % op)
    out = torch.empty_like(x) if out is None else out
    if out.numel() != x.numel() or out.dtype != x.dtype or not out.is_contiguous():
        raise ValueError("map2 %s: out must be a contiguous table like x" % op)
    f = lib().nt2h_map2_f32 if x.dtype == torch.float32 else lib().nt2h_map2_f64
    check(f(OPS2[op], x.numel(), ptr(x), ptr(y), ptr(out), policy, _s()), "map2 " + op)
    return out


def saxpy(a, x, y):
    check(lib().nt2h_saxpy_f32(x.numel(), ctypes.c_float(a), ptr(x), ptr(y), _s()), "saxpy")
    return y


def sum(x, out=None):
    out = torch.empty(1, dtype=torch.float64, device=x.device) if out is None else out
    check(lib().nt2h_sum_f32(x.numel(), ptr(x), ptr(out), _s()), "sum")
    return out


def dot(x, y, out=None):
    out = torch.empty(1, dtype=torch.float64, device=x.device) if out is None else out
    check(lib().nt2h_dot_f32(x.numel(), ptr(x), ptr(y), ptr(out), _s()), "dot")
    return out


def blackscholes(S, X, T, r, v, mode=BS_TEXTBOOK, policy=EXACT, call=None, put=None):
    call = torch.empty_like(S) if call is None else call
    if S.dtype == torch.float32:
        if put is None and mode == BS_TEXTBOOK:
            put = torch.empty_like(S)
        check(lib().nt2h_blackscholes_f32(S.numel(), ptr(S), ptr(X), ptr(T), ctypes.c_float(r),
                                          ctypes.c_float(v), ptr(call), ptr(put), mode, policy,
                                          _s()), "blackscholes_f32")
    else:
        put = torch.empty_like(S) if put is None else put
        check(lib().nt2h_blackscholes_f64(S.numel(), ptr(S), ptr(X), ptr(T), r, v, ptr(call),
                                          ptr(put), policy, _s()), "blackscholes_f64")
    return call, put


def sigmadelta(frames, N, counts=True, bkg=None, var=None, label=None, cnt=None):
    nf, npix = frames.shape
    bkg = frames[0].clone() if bkg is None else bkg
    var = torch.ones(npix, dtype=torch.uint8, device=frames.device) if var is None else var
    label = torch.zeros(npix, dtype=torch.uint8, device=frames.device) if label is None else label
    if cnt is None and counts:
        cnt = torch.zeros(nf, dtype=torch.int32, device=frames.device)
    check(lib().nt2h_sigmadelta_u8(npix, nf, ptr(frames), frames.stride(0), ptr(bkg), ptr(var),
                                   ptr(label), N, ptr(cnt), _s()), "sigmadelta")
    return bkg, var, label, cnt


def sigmadelta_step(frame, bkg, var, label, N):
    check(lib().nt2h_sigmadelta_step_u8(frame.numel(), ptr(frame), ptr(bkg), ptr(var), ptr(label),
                                        N, _s()), "sigmadelta_step")


def expr4(x, y, policy=EXACT):
    z = torch.empty_like(x)
    check(lib().nt2h_expr4_f32(x.numel(), ptr(x), ptr(y), ptr(z), policy, _s()), "expr4")
    return z


def expr4_sum(x, y, policy=EXACT, out=None):
    out = torch.empty(1, dtype=torch.float64, device=x.device) if out is None else out
    check(lib().nt2h_expr4_sum_f32(x.numel(), ptr(x), ptr(y), ptr(out), policy, _s()), "expr4_sum")
    return out


def _ext4(shape):
    """NT2 extent (column-major d0..d3) of a table given as a sequence of at
    most 4 sizes; trailing sizes are 1."""
    e = list(shape) + [1] * (4 - len(shape))
    if len(e) != 4:
        raise ValueError("at most 4 dimensions")
    arr = (ctypes.c_size_t * 4)(*e)
    return arr, e


def sum_dim(x, ext, dim, policy=EXACT, out=None):
    """NT2's sum(a, dim) of the device table x (flat, column-major, extent
    `ext`); returns the flat result of extent ext with ext[dim-1] = 1."""
    e4, e = _ext4(ext)
    if dim <= 4:
        e[dim - 1] = 1
    n = 1
    for v in e:
        n *= v
    out = torch.empty(n, dtype=x.dtype, device=x.device) if out is None else out
    f = lib().nt2h_sum_dim_f32 if x.dtype == torch.float32 else lib().nt2h_sum_dim_f64
    check(f(ptr(x), ctypes.cast(e4, ctypes.c_void_p), dim, ptr(out), policy, _s()), "sum_dim")
    return out


def cumsum_dim(x, ext, dim, policy=EXACT, out=None, op="cumsum"):
    """NT2's cumsum(a, dim) of the device table x (flat, column-major); op
    "cumprod", "cummin" or "cummax" runs that scan instead."""
    if op not in ("cumsum", "cumprod", "cummin", "cummax"):
        raise ValueError("unknown scan %r" % op)
    e4, _ = _ext4(ext)
    out = torch.empty_like(x) if out is None else out
    f = getattr(lib(), "nt2h_%s_dim_%s" % (op, "f32" if x.dtype == torch.float32 else "f64"))
    check(f(ptr(x), ctypes.cast(e4, ctypes.c_void_p), dim, ptr(out), policy, _s()), op + "_dim")
    return out


def rgb2grey(r, g, b, policy=EXACT, out=None):
    """grey = 0.3 r + 0.59 g + 0.11 b over planar float32 device tables."""
    out = torch.empty_like(r) if out is None else out
    check(lib().nt2h_rgb2grey_f32(r.numel(), ptr(r), ptr(g), ptr(b), ptr(out), policy, _s()), "rgb2grey")
    return out


def rgb2grey_u8(r, g, b, out=None):
    """grey = min(255, (77 r + 151 g + 28 b + 128) >> 8) over uint8 device planes."""
    out = torch.empty_like(r) if out is None else out
    check(lib().nt2h_rgb2grey_u8(r.numel(), ptr(r), ptr(g), ptr(b), ptr(out), _s()), "rgb2grey_u8")
    return out


def rgb2yuv(r, g, b, policy=EXACT):
    y, u, v = torch.empty_like(r), torch.empty_like(r), torch.empty_like(r)
    check(lib().nt2h_rgb2yuv_f32(r.numel(), ptr(r), ptr(g), ptr(b), ptr(y), ptr(u), ptr(v), policy, _s()),
          "rgb2yuv")
    return y, u, v


def linspace(lo, hi, n, dtype=torch.float32, policy=EXACT, device="cuda"):
    out = torch.empty(n, dtype=dtype, device=device)
    f = lib().nt2h_linspace_f32 if dtype == torch.float32 else lib().nt2h_linspace_f64
    check(f(n, lo, hi, ptr(out), policy, _s()), "linspace")
    return out


def mandelbrot(a, b, max_iter=256, policy=EXACT, out=None):
    """Escape counts (int32) of the points a + ib, NT2's pack-uniform loop."""
    out = torch.empty(a.numel(), dtype=torch.int32, device=a.device) if out is None else out
    check(lib().nt2h_mandelbrot_f32(a.numel(), ptr(a), ptr(b), ptr(out), int(max_iter), policy, _s()),
          "mandelbrot")
    return out
