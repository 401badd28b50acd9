"""Launch sequences for rocprofv3 runs (kernel traces and HBM byte counters).

    python3 -m nt2_amd.profile_workloads bs      # Black-Scholes f32 2^26, every policy, x3
    python3 -m nt2_amd.profile_workloads bs64    # Black-Scholes f64 2^26, accurate and exact, x3
    python3 -m nt2_amd.profile_workloads fold    # saxpy + dot 2^20 and 2^26
    python3 -m nt2_amd.profile_workloads sd      # sigma-delta 64 x 1920x1080
    python3 -m nt2_amd.profile_workloads batch   # 4096 uneven tables, persistent (256x32, 8/CU) vs plain
    python3 -m nt2_amd.profile_workloads batchp  # the persistent schedule alone
    python3 -m nt2_amd.profile_workloads dimwise # sum / cumsum of 8192 x 8192 f32, dims 1 and 2, both policies
    python3 -m nt2_amd.profile_workloads images  # rgb2grey / rgb2yuv 8192 x 8192, mandelbrot 1600 x 1600

Each run also launches a 1 GiB streaming copy: the calibration kernel that
turns FETCH_SIZE / WRITE_SIZE counter values into bytes.
"""
import sys

import numpy as np
import torch

from . import ops
from .native import ACCURATE, EXACT, FAST



def _copy():
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    ops.stream_copy(a, b)
    return a, b


def bs(reps=3):
    n = 1 << 26
    S = ops.fill_uniform(n, 5.0, 30.0, 33)
    X = ops.fill_uniform(n, 1.0, 100.0, 34)
    T = ops.fill_uniform(n, 0.25, 10.0, 35)
    c = torch.empty_like(S)
    p = torch.empty_like(S)
    for _ in range(reps):
        for pol in (ACCURATE, FAST, EXACT):
            ops.blackscholes(S, X, T, 0.02, 0.30, policy=pol, call=c, put=p)
        _copy()


def bs64(reps=3):
    n = 1 << 26
    S = ops.fill_uniform(n, 5.0, 30.0, 33, dtype=torch.float64)
    X = ops.fill_uniform(n, 1.0, 100.0, 34, dtype=torch.float64)
    T = ops.fill_uniform(n, 0.25, 10.0, 35, dtype=torch.float64)
    c = torch.empty_like(S)
    p = torch.empty_like(S)
    for _ in range(reps):
        for pol in (ACCURATE, EXACT):
            ops.blackscholes(S, X, T, 0.02, 0.30, policy=pol, call=c, put=p)
        _copy()


def fold(reps=3):
    for n in (1 << 20, 1 << 26):
        x = ops.fill_uniform(n, -1.0, 1.0, 1)
        y = ops.fill_uniform(n, -1.0, 1.0, 11)
        for _ in range(reps):
            ops.saxpy(2.5, x, y)
            ops.dot(x, y)
    _copy()


def sd(reps=3):
    frames = ops.fill_frames(1080, 1920, 64, seed=3)
    for _ in range(reps):
        ops.sigmadelta(frames, 2)
    _copy()


def batch(reps=2, count=4096, seed=4, plain=True):
    from .batch import Batch, new_sums, uneven_tables
    x, y, offs, lens = uneven_tables(count, seed=seed)
    sums = new_sums(count)
    with Batch(x, y, offs, lens, FAST) as b:
        b.set_config(256, 32, 8)  # the autotuner's choice for this workload on MI355X
        for _ in range(reps):
            b.run(sums)
            if plain:
                b.run_plain(sums)
        torch.cuda.synchronize()
    _copy()


def batchp():
    """the persistent schedule alone (counter passes: 4096 short plain
    launches per run add nothing to its figures)"""
    batch(plain=False)


def dimwise(reps=2):
    n = 8192
    x = ops.fill_uniform(n * n, -1.0, 1.0, 81)
    out = torch.empty_like(x)
    for _ in range(reps):
        for dim in (1, 2):
            for pol in (EXACT, FAST):
                ops.sum_dim(x, (n, n), dim, pol)
                ops.cumsum_dim(x, (n, n), dim, pol, out=out)
    _copy()


def images(reps=3):
    n = 8192 * 8192
    r, g, b = (ops.fill_uniform(n, 0.0, 255.0, 90 + k) for k in range(3))
    grey = torch.empty_like(r)
    y, u, v = torch.empty_like(r), torch.empty_like(r), torch.empty_like(r)
    m = 1600
    la = ops.linspace(-1.5, 1.5, m)
    A, B = la.repeat_interleave(m), la.repeat(m)
    C = torch.empty(m * m, dtype=torch.int32, device="cuda")
    from .native import check, lib, ptr
    r8, g8, b8 = (torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda") for _ in range(3))
    grey8 = torch.empty_like(r8)
    for _ in range(reps):
        ops.rgb2grey_u8(r8, g8, b8, out=grey8)
        for pol in (EXACT, FAST):
            ops.rgb2grey(r, g, b, pol, out=grey)
            check(lib().nt2h_rgb2yuv_f32(n, ptr(r), ptr(g), ptr(b), ptr(y), ptr(u), ptr(v), pol, None), "rgb2yuv")
            ops.mandelbrot(A, B, 256, pol, out=C)
    _copy()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "bs"
    {"bs": bs, "bs64": bs64, "fold": fold, "sd": sd, "batch": batch, "batchp": batchp, "dimwise": dimwise,
     "images": images}[what]()
    torch.cuda.synchronize()
    print("profile_workloads %s done" % what, flush=True)
