"""Launch sequences for rocprofv3 runs (kernel traces and HBM byte counters).

    python3 -m nt2_amd.profile_workloads bs      # Black-Scholes f32 2^26, every policy, x3
    python3 -m nt2_amd.profile_workloads fold    # saxpy + dot 2^20 and 2^26
    python3 -m nt2_amd.profile_workloads sd      # sigma-delta 64 x 1920x1080
    python3 -m nt2_amd.profile_workloads batch   # 4096 uneven tables, persistent vs plain

Each run also launches a 1 GiB streaming copy: the calibration kernel that
turns FETCH_SIZE / WRITE_SIZE counter values into bytes.
"""
import sys

import numpy as np
import torch

from . import ops
from .native import EXACT, FAST

ACCURATE = 2


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


def batch(reps=2, count=4096, seed=4):
    import ctypes
    from .native import check, lib, ptr, stream_handle
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(1e6), count))).astype(np.int64)
    total = int(lens.sum())
    x = ops.fill_normal(total, 41)
    y = ops.fill_uniform(total, -1e4, 1e4, 42)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]])
    # 16-byte aligned sub-tables: round each offset up to a multiple of 4 floats
    offs = (offs + 3) // 4 * 4
    need = int(offs[-1] + lens[-1])
    if need > total:
        x = torch.cat([x, ops.fill_normal(need - total, 43)])
        y = torch.cat([y, ops.fill_uniform(need - total, -1e4, 1e4, 44)])
    XP = (ctypes.c_void_p * count)(*[x.data_ptr() + 4 * int(o) for o in offs])
    YP = (ctypes.c_void_p * count)(*[y.data_ptr() + 4 * int(o) for o in offs])
    LN = (ctypes.c_size_t * count)(*[int(n) for n in lens])
    b = ctypes.c_void_p()
    check(lib().nt2h_batch_create(ctypes.byref(b), count, XP, YP, LN, FAST), "batch_create")
    sums = torch.zeros(count, dtype=torch.float64, device="cuda")
    for _ in range(reps):
        check(lib().nt2h_batch_run(b, ptr(sums), stream_handle()), "batch_run")
        check(lib().nt2h_batch_run_plain(b, ptr(sums), stream_handle()), "batch_plain")
    torch.cuda.synchronize()
    lib().nt2h_batch_destroy(b)
    _copy()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "bs"
    {"bs": bs, "fold": fold, "sd": sd, "batch": batch}[what]()
    torch.cuda.synchronize()
    print("profile_workloads %s done" % what, flush=True)
