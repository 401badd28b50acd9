"""nt2_amd — MI355X (gfx950) architecture site for NT2's table-expression path.

The product is the native library libnt2h.so (C ABI, include/nt2h.h) and the
C++ front end under nt2_amd/include/nt2/.  This module is a thin ctypes
binding of the C ABI used by the tests and the benchmark; device memory and
streams come from torch tensors (plumbing only).  Importing it does not
touch the GPU; calling anything without the built library raises.
"""
from . import native  # noqa: F401
from .native import (  # noqa: F401
    EXACT, FAST, ACCURATE, OPS, BS_TEXTBOOK, BS_DEMO, Nt2hError, lib, lib_path,
)
