"""nt2h_dot_f32_noalloc (the fold with its partials in a static device
array) against nt2h_dot_f32, through the C ABI: bit-identical results on two
calls in a row, the second on a smaller input, and within 2^-40 of the exact
sum of the float products."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import _oracle as O  # noqa: E402
from _gpu import say  # noqa: E402
from nt2_amd import ops  # noqa: E402
from nt2_amd.native import check, lib, ptr, stream_handle  # noqa: E402

pytestmark = pytest.mark.gpu


def _dot_both(x, y, n):
    a = torch.full((1,), 7.0, dtype=torch.float64, device="cuda")
    b = torch.full((1,), -7.0, dtype=torch.float64, device="cuda")
    check(lib().nt2h_dot_f32(n, ptr(x), ptr(y), ptr(a), stream_handle()), "dot")
    check(lib().nt2h_dot_f32_noalloc(n, ptr(x), ptr(y), ptr(b), stream_handle()), "dot_noalloc")
    torch.cuda.synchronize()
    return a.item(), b.item()


@pytest.mark.parametrize("n1,n2", [((1 << 24) + 3, 1000), (1 << 20, 5)])
def test_noalloc_dot_equals_dot(n1, n2):
    x = ops.fill_uniform(n1, -1.0, 1.0, 71)
    y = ops.fill_uniform(n1, -1.0, 1.0, 72)
    for n in (n1, n2):  # the second call on a smaller input reuses the array
        say("dot_noalloc n=%d" % n)
        a, b = _dot_both(x, y, n)
        assert np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64), (n, a, b)
        p = (x[:n].cpu().numpy() * y[:n].cpu().numpy()).astype(np.float32)
        ex, mag = O.sum_f32_exact(p), O.sum_f32_exact(np.abs(p))
        assert abs(b - ex) <= 2 ** -40 * mag + 1e-30


def test_noalloc_dot_unaligned_and_empty():
    x = ops.fill_uniform(4099, -1.0, 1.0, 73)
    y = ops.fill_uniform(4099, -1.0, 1.0, 74)
    say("dot_noalloc unaligned")
    xa, ya = x[1:], y[3:]
    a = torch.empty(1, dtype=torch.float64, device="cuda")
    b = torch.empty(1, dtype=torch.float64, device="cuda")
    check(lib().nt2h_dot_f32(4096, ptr(xa), ptr(ya), ptr(a), stream_handle()), "dot")
    check(lib().nt2h_dot_f32_noalloc(4096, ptr(xa), ptr(ya), ptr(b), stream_handle()), "dot_noalloc")
    torch.cuda.synchronize()
    assert a.item() == b.item()
    say("dot_noalloc n=0")
    b.fill_(5.0)
    check(lib().nt2h_dot_f32_noalloc(0, None, None, ptr(b), stream_handle()), "dot_noalloc 0")
    torch.cuda.synchronize()
    assert b.item() == 0.0
    assert lib().nt2h_dot_f32_noalloc(16, None, ptr(y), ptr(b), stream_handle()) == -1
    assert lib().nt2h_dot_f32_noalloc(16, ptr(x), ptr(y), None, stream_handle()) == -1
