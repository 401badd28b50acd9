"""Serving-engine scheduler on the host back end (CPU, no GPU): admission in
arrival order against the budget, malformed and oversized requests, forced
allocation failures, cancellation, results and statistics."""
import ctypes
import time

import numpy as np
import pytest
from scipy.special import ndtr

from nt2_amd import native
from nt2_amd.serve import BACKEND_HOST, EAGAIN, F32, F64, Server

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(native.lib_path()),
                                reason="libnt2h.so not built")

EINVAL, EBUDGET, ENOMEM, ECANCELLED, ESTATE = -1, -4, -3, -6, -7


def _tables(n, dtype, seed):
    rng = np.random.default_rng(seed)
    S = rng.uniform(5, 30, n).astype(dtype)
    X = rng.uniform(1, 100, n).astype(dtype)
    T = rng.uniform(0.25, 10, n).astype(dtype)
    return S, X, T


def _ref(S, X, T, r, v):
    S, X, T = (a.astype(np.float64) for a in (S, X, T))
    vst = v * np.sqrt(T)
    d1 = (np.log(S / X) + (r + v * v / 2) * T) / vst
    d2 = d1 - vst
    e = X * np.exp(-r * T)
    return S * ndtr(d1) - e * ndtr(d2), e * ndtr(-d2) - S * ndtr(-d1)


def _p(a):
    return a.ctypes.data


def test_results_and_stats():
    keep = []
    with Server(1 << 30, max_inflight=2, max_batch=8, backend=BACKEND_HOST) as srv:
        ids = []
        for k in range(20):
            dt = np.float64 if k % 5 == 0 else np.float32
            n = 100 + 37 * k
            S, X, T = _tables(n, dt, k)
            keep.append((S, X, T))
            rc, rid = srv.submit(F64 if dt is np.float64 else F32, n, _p(S), _p(X), _p(T), 0.02, 0.30)
            assert rc == 0
            ids.append(rid)
        for k, rid in enumerate(ids):
            assert srv.wait(rid, 30) == 0
            rc, c, p, lat = srv.result(rid)
            assert rc == 0 and lat > 0
            S, X, T = keep[k]
            ct = ctypes.c_double if S.dtype == np.float64 else ctypes.c_float
            call = np.ctypeslib.as_array((ct * S.size).from_address(c))
            put = np.ctypeslib.as_array((ct * S.size).from_address(p))
            rc_, rp_ = _ref(S, X, T, 0.02, 0.30)
            tol = 1e-9 if S.dtype == np.float64 else 2e-3
            assert np.max(np.abs(call - rc_)) < tol and np.max(np.abs(put - rp_)) < tol
            assert srv.release(rid) == 0
        st = srv.stats()
        assert st["completed"] == 20 and st["failed"] == 0
        assert st["p50_us"] <= st["p99_us"] <= st["max_us"]
        assert st["used_bytes"] == 0


def test_malformed_and_oversized():
    S, X, T = _tables(16, np.float32, 1)
    with Server(1000, backend=BACKEND_HOST) as srv:
        assert srv.submit(F32, 0, _p(S), _p(X), _p(T), 0.02, 0.3)[0] == EINVAL
        assert srv.submit(F32, 16, None, _p(X), _p(T), 0.02, 0.3)[0] == EINVAL
        assert srv.submit(7, 16, _p(S), _p(X), _p(T), 0.02, 0.3)[0] == EINVAL
        assert srv.submit(F32, 16, _p(S), _p(X), _p(T), 0.02, -0.3)[0] == EINVAL
        # 5 tables * 4 B * 100 options = 2000 B > budget 1000 B
        S2, X2, T2 = _tables(100, np.float32, 2)
        assert srv.submit(F32, 100, _p(S2), _p(X2), _p(T2), 0.02, 0.3)[0] == EBUDGET
        rc, rid = srv.submit(F32, 16, _p(S), _p(X), _p(T), 0.02, 0.3)
        assert rc == 0 and srv.wait(rid) == 0
        assert srv.release(rid) == 0
        assert srv.poll(rid) == EINVAL  # released requests are forgotten


def test_budget_admits_in_arrival_order():
    """Budget of 3 requests: the 4th waits until one is released, and requests
    are admitted in arrival order."""
    n = 1000
    per = 5 * n * 4
    tabs = [_tables(n, np.float32, k) for k in range(6)]
    with Server(3 * per, max_inflight=4, max_batch=64, backend=BACKEND_HOST) as srv:
        ids = [srv.submit(F32, n, _p(S), _p(X), _p(T), 0.02, 0.3)[1] for (S, X, T) in tabs]
        for rid in ids[:3]:
            assert srv.wait(rid, 30) == 0
        time.sleep(0.05)
        assert srv.poll(ids[3]) == EAGAIN  # held back by the budget
        assert srv.stats()["used_bytes"] == 3 * per
        assert srv.release(ids[0]) == 0
        assert srv.wait(ids[3], 30) == 0
        assert srv.poll(ids[4]) == EAGAIN
        for rid in ids[1:4]:
            srv.release(rid)
        for rid in ids[4:]:
            assert srv.wait(rid, 30) == 0
            srv.release(rid)
        assert srv.stats()["peak_bytes"] <= 3 * per


def test_cancel_before_launch():
    n = 1000
    per = 5 * n * 4
    tabs = [_tables(n, np.float32, k) for k in range(3)]
    with Server(per, max_inflight=1, backend=BACKEND_HOST) as srv:
        a = srv.submit(F32, n, *(_p(t) for t in tabs[0]), 0.02, 0.3)[1]
        b = srv.submit(F32, n, *(_p(t) for t in tabs[1]), 0.02, 0.3)[1]
        assert srv.cancel(b) == 0           # still waiting for budget
        assert srv.wait(a) == 0
        assert srv.cancel(a) == ESTATE      # already finished
        assert srv.poll(b) == ECANCELLED
        srv.release(a)
        c = srv.submit(F32, n, *(_p(t) for t in tabs[2]), 0.02, 0.3)[1]
        assert srv.wait(c) == 0
        assert srv.stats()["cancelled"] == 1


def test_allocation_failure_is_isolated():
    """Every 3rd output allocation fails: those requests fail with ENOMEM,
    the others complete, and the budget is returned."""
    tabs = [_tables(64, np.float32, k) for k in range(9)]
    with Server(1 << 24, max_inflight=1, max_batch=1, backend=BACKEND_HOST, fail_alloc_every=3) as srv:
        ids = [srv.submit(F32, 64, *(_p(t) for t in tb), 0.02, 0.3)[1] for tb in tabs]
        codes = [srv.wait(r, 30) for r in ids]
        assert codes.count(ENOMEM) == 3 and codes.count(0) == 6
        for r in ids:
            srv.release(r)
        st = srv.stats()
        assert st["failed"] == 3 and st["completed"] == 6 and st["used_bytes"] == 0


def test_launch_failure_returns_budget_and_outputs():
    """Every 3rd launched batch fails after launch (a test hook standing in for
    a device error): its requests fail with EHIP and give their bytes back at
    once, so that a budget of two requests keeps admitting the rest; nothing
    stays charged after the releases."""
    n = 64
    tabs = [_tables(n, np.float32, k) for k in range(12)]
    budget = 2 * 5 * n * 4
    EHIP = -2
    with Server(budget, max_inflight=1, max_batch=1, backend=BACKEND_HOST, fail_launch_every=3) as srv:
        ids = [srv.submit(F32, n, *(_p(t) for t in tb), 0.02, 0.3)[1] for tb in tabs]
        codes = []
        for r in ids:  # completed results are released at once, failed ones only at the end
            codes.append(srv.wait(r, 30))
            if codes[-1] == 0:
                srv.release(r)
        assert codes.count(EHIP) == 4 and codes.count(0) == 8, codes
        assert srv.result(ids[2])[0] == EHIP
        assert srv.stats()["used_bytes"] == 0  # the failed requests hold no budget
        for r, c in zip(ids, codes):
            if c != 0:
                srv.release(r)
        st = srv.stats()
        assert st["failed"] == 4 and st["completed"] == 8 and st["used_bytes"] == 0
    # batched: one failing launch fails all of its requests
    with Server(1 << 24, max_inflight=1, max_batch=64, backend=BACKEND_HOST, fail_launch_every=1) as srv:
        ids = [srv.submit(F32, n, *(_p(t) for t in tb), 0.02, 0.3)[1] for tb in tabs]
        assert all(srv.wait(r, 30) == EHIP for r in ids)
        for r in ids:
            srv.release(r)
        assert srv.stats()["used_bytes"] == 0


def test_plain_schedule_one_request_per_launch():
    tabs = [_tables(200, np.float32, k) for k in range(10)]
    with Server(1 << 24, max_inflight=1, max_batch=1, backend=BACKEND_HOST) as srv:
        ids = [srv.submit(F32, 200, *(_p(t) for t in tb), 0.02, 0.3)[1] for tb in tabs]
        assert srv.drain(30) == 0
        st = srv.stats()
        assert st["launches"] == 10 and st["mean_batch"] == 1.0
        for r in ids:
            srv.release(r)


def test_trace_replay_host():
    tabs = [_tables(256, np.float32, k) for k in range(50)]
    from nt2_amd.serve import Request
    items = [Request(F32, 256, _p(S), _p(X), _p(T), 0.02, 0.3, None, None, 0) for (S, X, T) in tabs]
    arrivals = list(np.cumsum(np.random.default_rng(0).exponential(2e-4, 50)))
    with Server(1 << 24, max_inflight=2, backend=BACKEND_HOST) as srv:
        assert srv.run_trace(items, arrivals, 60) == 0
        st = srv.stats()
        assert st["completed"] == 50 and st["used_bytes"] == 0 and st["options"] == 50 * 256


def test_trace_replay_count_bounded():
    """A replay never reads past the prepared requests or arrival times."""
    tabs = [_tables(64, np.float32, k) for k in range(4)]
    from nt2_amd.serve import Request
    items = (Request * 4)(*[Request(F32, 64, _p(S), _p(X), _p(T), 0.02, 0.3, None, None, 0) for (S, X, T) in tabs])
    with Server(1 << 24, max_inflight=2, backend=BACKEND_HOST) as srv:
        with pytest.raises(ValueError):
            srv.run_trace(items, np.zeros(8), 10, count=5)
        with pytest.raises(ValueError):
            srv.run_trace(items, np.zeros(3), 10, count=4)
        assert srv.run_trace(items, np.zeros(4), 10, count=4) == 0
        assert srv.stats()["completed"] == 4


def test_concurrent_clients_stress():
    """tests/cpp/server_stress: four client threads submitting, polling,
    waiting, cancelling and releasing concurrently against the host back end
    (forced allocation and launch failures); every request ends and the
    budget returns to zero."""
    import json
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "server_stress")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/bin/server_stress not built")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    r = json.loads(p.stdout.strip().splitlines()[-1])
    assert r["used_bytes"] == 0 and r["done"] > 0 and r["failed"] > 0 and r["cancelled"] > 0
