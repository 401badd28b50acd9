"""ctypes binding of the serving engine (nt2h_server_* in include/nt2h.h)."""
import ctypes

from .native import check, lib

F32, F64 = 0, 1
BACKEND_HIP, BACKEND_HOST = 0, 1
EAGAIN = -5


class Config(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_size_t), ("max_inflight", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("policy", ctypes.c_int), ("backend", ctypes.c_int),
                ("fail_alloc_every", ctypes.c_int), ("fail_launch_every", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int), ("n", ctypes.c_size_t), ("S", ctypes.c_void_p),
                ("X", ctypes.c_void_p), ("T", ctypes.c_void_p), ("r", ctypes.c_double),
                ("v", ctypes.c_double), ("call", ctypes.c_void_p), ("put", ctypes.c_void_p),
                ("t_arrival_ns", ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("options", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("mean_batch", ctypes.c_double),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("max_us", ctypes.c_double),
                ("span_s", ctypes.c_double), ("options_per_s", ctypes.c_double),
                ("used_bytes", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _bind():
    L = lib()
    P, I, D, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64
    sig = {
        "nt2h_server_create": ([ctypes.POINTER(Config), P], I),
        "nt2h_server_submit": ([P, ctypes.POINTER(Request), ctypes.POINTER(U64)], I),
        "nt2h_server_poll": ([P, U64], I),
        "nt2h_server_wait": ([P, U64, D], I),
        "nt2h_server_result": ([P, U64, P, P, P], I),
        "nt2h_server_cancel": ([P, U64], I),
        "nt2h_server_release": ([P, U64], I),
        "nt2h_server_drain": ([P, D], I),
        "nt2h_server_get_stats": ([P, ctypes.POINTER(Stats)], I),
        "nt2h_server_reset_stats": ([P], I),
        "nt2h_server_run_trace": ([P, I, ctypes.POINTER(Request), ctypes.POINTER(D), D], I),
        "nt2h_server_destroy": ([P], I),
    }
    for name, (a, r) in sig.items():
        f = getattr(L, name)
        f.argtypes = a
        f.restype = r
    return L


class Server:
    """Owns one nt2h_server."""

    def __init__(self, budget_bytes, max_inflight=4, max_batch=4096, policy=1,
                 backend=BACKEND_HIP, fail_alloc_every=0, fail_launch_every=0):
        self.L = _bind()
        cfg = Config(budget_bytes, max_inflight, max_batch, policy, backend, fail_alloc_every,
                     fail_launch_every)
        self.h = ctypes.c_void_p()
        check(self.L.nt2h_server_create(ctypes.byref(cfg), ctypes.byref(self.h)), "server_create")

    def submit(self, dtype, n, S, X, T, r, v, call=None, put=None, t_arrival_ns=0):
        q = Request(dtype, n, S, X, T, r, v, call, put, t_arrival_ns)
        rid = ctypes.c_uint64()
        rc = self.L.nt2h_server_submit(self.h, ctypes.byref(q), ctypes.byref(rid))
        return rc, rid.value

    def poll(self, rid):
        return self.L.nt2h_server_poll(self.h, rid)

    def wait(self, rid, timeout=30.0):
        return self.L.nt2h_server_wait(self.h, rid, timeout)

    def result(self, rid):
        c, p, lat = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int64()
        rc = self.L.nt2h_server_result(self.h, rid, ctypes.byref(c), ctypes.byref(p), ctypes.byref(lat))
        return rc, c.value, p.value, lat.value

    def cancel(self, rid):
        return self.L.nt2h_server_cancel(self.h, rid)

    def release(self, rid):
        return self.L.nt2h_server_release(self.h, rid)

    def drain(self, timeout=60.0):
        return self.L.nt2h_server_drain(self.h, timeout)

    def stats(self):
        st = Stats()
        check(self.L.nt2h_server_get_stats(self.h, ctypes.byref(st)), "server_stats")
        return st.as_dict()

    def reset_stats(self):
        check(self.L.nt2h_server_reset_stats(self.h), "server_reset_stats")

    def run_trace(self, items, arrivals, timeout=120.0, count=None):
        """Replay items[:count] (a list of Request or a ctypes Request array)
        at the given arrival times (seconds from the start)."""
        n = len(items) if count is None else count
        if n < 0 or n > len(items) or n > len(arrivals):
            raise ValueError("run_trace: count %d exceeds the %d items or %d arrival times"
                             % (n, len(items), len(arrivals)))
        arr = items if isinstance(items, ctypes.Array) else (Request * n)(*items)
        ts = (ctypes.c_double * n)(*[float(a) for a in arrivals[:n]])
        return self.L.nt2h_server_run_trace(self.h, n, arr, ts, timeout)

    def close(self):
        if self.h:
            self.L.nt2h_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
