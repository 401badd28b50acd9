"""The C-ABI library loads and exports every entry point include/nt2h.h
declares (CPU only: nothing here touches a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nt2h.h")
LIB = os.path.join(ROOT, "nt2_amd", "libnt2h.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(nt2h_\w+)\s*\(", text, re.M)))


def test_header_declares_the_path():
    names = declared()
    for must in ("nt2h_blackscholes_f32", "nt2h_blackscholes_f64", "nt2h_saxpy_f32", "nt2h_dot_f32",
                 "nt2h_sum_f32", "nt2h_sigmadelta_u8", "nt2h_map_f32", "nt2h_batch_create",
                 "nt2h_batch_run", "nt2h_server_create", "nt2h_server_submit"):
        assert must in names
    assert len(names) >= 40


def test_op_codes_agree_with_the_bindings():
    """nt2h_map_f32/f64 op codes: the header's NT2H_OP_* values are the ones the
    Python bindings and the oracle bindings send."""
    import sys
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from nt2_amd.native import OPS
    import _oracle
    ops = {m.group(1).lower(): int(m.group(2))
           for m in re.finditer(r"#define NT2H_OP_(\w+) (\d+)", open(HEADER).read())}
    assert ops == OPS == _oracle.OPS


@pytest.mark.skipif(not os.path.exists(LIB), reason="libnt2h.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.nt2h_abi_version.restype = ctypes.c_int
    assert lib.nt2h_abi_version() >= 1
    lib.nt2h_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.nt2h_last_error(), (bytes, type(None)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libnt2h.so not built (run __graft_entry__.build())")
def test_argument_errors_are_codes_not_crashes():
    """Null pointers and bad codes are rejected before any device work."""
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.nt2h_blackscholes_f32.argtypes = [ctypes.c_size_t, P, P, P, ctypes.c_float, ctypes.c_float, P, P,
                                          ctypes.c_int, ctypes.c_int, P]
    assert lib.nt2h_blackscholes_f32(16, None, None, None, 0.02, 0.3, None, None, 0, 0, None) == -1
    lib.nt2h_map_f32.argtypes = [ctypes.c_int, ctypes.c_size_t, P, P, ctypes.c_int, P]
    assert lib.nt2h_map_f32(99, 16, P(16), P(16), 0, None) == -1
    lib.nt2h_last_error.restype = ctypes.c_char_p
    assert lib.nt2h_last_error()
