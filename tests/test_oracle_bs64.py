"""The quad-precision Black-Scholes reference (oracle_bs_quad_f64in), the
yardstick of the accurate f64 policy, pinned by mpmath prices at 60 digits
(tests/golden/bs64_golden.npz: the benchmark distribution, other rates and
volatilities, short maturities with tails down to ~1e-250, near the money)
and by put-call parity."""
import os

import numpy as np

import _oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    return np.load(os.path.join(G, "bs64_golden.npz"))


def test_quad_reference_matches_mpmath():
    g = _golden()
    for (r, v) in sorted(set(zip(g["r"], g["v"]))):
        m = (g["r"] == r) & (g["v"] == v)
        ch, cl, ph, pl = O.bs_quad(g["S"][m], g["X"][m], g["T"][m], float(r), float(v))
        for hi, lo, wh, wl in ((ch, cl, g["call_hi"][m], g["call_lo"][m]), (ph, pl, g["put_hi"][m], g["put_lo"][m])):
            diff = np.abs((hi - wh) + (lo - wl))
            scale = np.abs(wh) + 1e-300
            # 113 bits less the cancellation of the two tails (up to ~2^14 for the
            # shortest maturities)
            assert np.all(diff <= 2.0 ** -90 * scale), (r, v, float(np.max(diff / scale)))


def test_quad_reference_put_call_parity_and_signs():
    rng = np.random.default_rng(12)
    n = 20000
    S, X, T = rng.uniform(5, 30, n), rng.uniform(1, 100, n), rng.uniform(0.01, 10, n)
    r, v = 0.02, 0.30
    ch, cl, ph, pl = O.bs_quad(S, X, T, r, v)
    F = X * np.exp(-r * T)
    assert np.all(ch >= 0) and np.all(ph >= 0)
    assert np.allclose((ch - ph) + (cl - pl), S - F, rtol=0, atol=1e-12 * (S + X))
    # a rising strike lowers the call and raises the put
    k = np.argsort(X[:200])
    S2, T2 = np.full(200, 17.0), np.full(200, 2.0)
    c2, _, p2, _ = O.bs_quad(S2, X[:200][k], T2, r, v)
    assert np.all(np.diff(c2) <= 0) and np.all(np.diff(p2) >= 0)


def test_fp64_textbook_loses_deep_tails():
    """Why the f64 yardstick is quad: the textbook formula evaluated in double
    is thousands of ulp from the mpmath price on deep out-of-the-money
    options (the two tails cancel)."""
    g = _golden()
    m = (g["T"] < 0.25) & (g["call_hi"] > 0) & (g["call_hi"] < 1e-30)
    assert m.sum() > 5
    c, _ = O.bs_fp64ref(g["S"][m], g["X"][m], g["T"][m], 0.02, 0.30)
    rel = np.abs(c - g["call_hi"][m]) / g["call_hi"][m]
    assert rel.max() > 1e-13
