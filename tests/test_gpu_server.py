"""Serving engine on the GPU: results are bit-identical alone and in a mix
of concurrently batched requests; a short Poisson trace completes with
latency statistics; the plain schedule gives the same bits."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from _gpu import say  # noqa: E402
from nt2_amd import ACCURATE, EXACT, FAST, ops  # noqa: E402
from nt2_amd.serve import F32, F64, Request, Server  # noqa: E402


def _req_tables(n, f64, seed):
    dt = torch.float64 if f64 else torch.float32
    S = ops.fill_uniform(n, 5.0, 30.0, seed * 3 + 1, dtype=dt)
    X = ops.fill_uniform(n, 1.0, 100.0, seed * 3 + 2, dtype=dt)
    T = ops.fill_uniform(n, 0.25, 10.0, seed * 3 + 3, dtype=dt)
    return S, X, T


def _outputs(srv, rid, n, f64):
    rc, c, p, lat = srv.result(rid)
    assert rc == 0
    dt = torch.float64 if f64 else torch.float32
    # copy the server-owned device tables through torch
    out_c = torch.empty(n, dtype=dt, device="cuda")
    out_p = torch.empty(n, dtype=dt, device="cuda")
    from nt2_amd.native import lib, check
    es = 8 if f64 else 4
    check(lib().nt2h_memcpy(ctypes.c_void_p(out_c.data_ptr()), ctypes.c_void_p(c), n * es, 3, None), "copy")
    check(lib().nt2h_memcpy(ctypes.c_void_p(out_p.data_ptr()), ctypes.c_void_p(p), n * es, 3, None), "copy")
    torch.cuda.synchronize()
    return out_c.cpu().numpy(), out_p.cpu().numpy()


@pytest.mark.parametrize("policy", [EXACT, FAST, ACCURATE])
def test_alone_equals_mixed(policy):
    """Under NT2H_ACCURATE the f32 and f64 requests use the accurate
    evaluations; under every policy each result is bit-identical alone, in the
    mix and from the standalone kernel."""
    rng = np.random.default_rng(5)
    sizes = [int(2 ** rng.uniform(10, 18)) for _ in range(40)]
    kinds = [bool(rng.random() < 0.2) for _ in sizes]
    tabs = [_req_tables(n, f, k) for k, (n, f) in enumerate(zip(sizes, kinds))]
    torch.cuda.synchronize()
    say("server mix policy=%d requests=%d" % (policy, len(sizes)))
    with Server(64 << 30, max_inflight=4, max_batch=4096, policy=policy) as srv:
        ids = []
        for (S, X, T), n, f in zip(tabs, sizes, kinds):
            rc, rid = srv.submit(F64 if f else F32, n, S.data_ptr(), X.data_ptr(), T.data_ptr(), 0.02, 0.30)
            assert rc == 0
            ids.append(rid)
        assert srv.drain(60) == 0
        mixed = [_outputs(srv, rid, n, f) for rid, n, f in zip(ids, sizes, kinds)]
        st = srv.stats()
        assert st["completed"] == len(sizes)
        for rid in ids:
            srv.release(rid)
    # each request alone, on a fresh plain server (one launch per request)
    with Server(64 << 30, max_inflight=1, max_batch=1, policy=policy) as srv:
        for k, ((S, X, T), n, f) in enumerate(zip(tabs, sizes, kinds)):
            rc, rid = srv.submit(F64 if f else F32, n, S.data_ptr(), X.data_ptr(), T.data_ptr(), 0.02, 0.30)
            assert rc == 0 and srv.wait(rid, 30) == 0
            c, p = _outputs(srv, rid, n, f)
            assert np.array_equal(c.view(np.uint8), mixed[k][0].view(np.uint8)), k
            assert np.array_equal(p.view(np.uint8), mixed[k][1].view(np.uint8)), k
            # and equal to the standalone kernel
            wc, wp = ops.blackscholes(S, X, T, 0.02, 0.30, policy=policy)
            assert np.array_equal(wc.cpu().numpy().view(np.uint8), c.view(np.uint8)), k
            assert np.array_equal(wp.cpu().numpy().view(np.uint8), p.view(np.uint8)), k
            srv.release(rid)


def test_poisson_trace_gpu():
    rng = np.random.default_rng(6)
    pool = 1 << 22
    S32, X32, T32 = _req_tables(pool, False, 50)
    S64, X64, T64 = _req_tables(pool, True, 51)
    torch.cuda.synchronize()
    count = 2000
    items, arr, t = [], [], 0.0
    for k in range(count):
        n = int(2 ** rng.uniform(10, 22))
        f = rng.random() < 0.2
        off = int(rng.integers(0, pool - n + 1))
        es = 8 if f else 4
        base = (S64, X64, T64) if f else (S32, X32, T32)
        items.append(Request(F64 if f else F32, n, *(b.data_ptr() + off * es for b in base),
                             0.02, 0.30, None, None, 0))
        t += rng.exponential(20e-6)
        arr.append(t)
    say("server trace count=%d" % count)
    with Server(64 << 30, max_inflight=4, max_batch=4096, policy=FAST) as srv:
        assert srv.run_trace(items, arr, 120) == 0
        st = srv.stats()
        assert st["completed"] == count and st["failed"] == 0 and st["used_bytes"] == 0
        assert 0 < st["p50_us"] <= st["p99_us"]
