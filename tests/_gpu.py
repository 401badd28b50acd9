"""Helpers for the GPU parity tests."""
import sys

import numpy as np


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    """Bit-exact equality where NaN == NaN (payloads differ between NT2's
    all-ones NaN constant and the hardware default)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        ia = a.view(np.uint32 if a.itemsize == 4 else np.uint64)
        ib = b.view(np.uint32 if b.itemsize == 4 else np.uint64)
        nan = np.isnan(a) & np.isnan(b)
        return bool(np.all((ia == ib) | nan))
    return bool(np.array_equal(a, b))


def first_diff(a, b, x=None):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.nonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))[0] if a.dtype.kind == "f" else np.nonzero(a != b)[0]
    if d.size == 0:
        return None
    i = d[0]
    return dict(index=int(i), count=int(d.size), got=a[i], want=b[i], x=None if x is None else x[i])


def splitmix_uniform(n, lo, hi, seed, dtype):
    """numpy restatement of the counter-based SplitMix64 table generator."""
    with np.errstate(over="ignore"):
        i = np.arange(n, dtype=np.uint64)
        z = np.uint64(seed) + (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (lo + (hi - lo) * u).astype(dtype)
