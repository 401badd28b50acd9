"""Compile-time behaviour of the C++ front end (nt2/table.hpp): the forms NT2
accepts compile, the forms this implementation does not provide stop the
compilation with a message naming the reason.  hipcc -fsyntax-only, no GPU."""
import os
import subprocess

import pytest

from nt2_amd import build as B

HIPCC = B.HIPCC
if not os.path.exists(HIPCC):
    pytest.skip("hipcc not available", allow_module_level=True)

PRELUDE = "#include <nt2/table.hpp>\nusing namespace nt2;\nint main() {\n"


def compile_snippet(tmp_path, body):
    src = tmp_path / "snippet.cpp"
    src.write_text(PRELUDE + body + "\nreturn 0;\n}\n")
    p = subprocess.run([HIPCC, "-x", "hip", "-std=c++17", "--offload-arch=gfx950", "-fsyntax-only",
                        "-I" + os.path.join(B.HERE, "include"), "-I" + os.path.join(B.ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=120)
    return p.returncode, p.stderr


ACCEPTED = [
    # NT2 settings: modules/core/sdk/include/nt2/core/settings/forward/*.hpp
    "table<float, _2D> a(of_size(3, 4)); a(3, 4) = 1.0f;",
    "table<double, C_index_> a(of_size(5)); a(0) = 1.0;",
    "table<float, settings(C_index_, _2D)> a(of_size(2, 2)); a(1, 1) = 2.0f;",
    "table<float, settings(matlab_order_, matlab_index_, owned_, aligned_, rectangular_, dynamic_)> a;",
    "table<float, _1D> a = reshape(_(1.0f, 4.0f), of_size(4)); table<float> b = a * 2.0f; float s = globalsum(a + b); (void)s;",
    "table<float, _3D> a(of_size(2, 2, 2)); table<float, _3D> b = sqrt(a); (void)b;",
    "table<int32_t> a(of_size(4)); int32_t s = globalsum(a); (void)s;",
    "table<double> a(of_size(4)); table<double> b = pow(a, 3); (void)b;",
]

REJECTED = [
    ("table<float, row_major_> a;", "only column-major, owned, dense storage"),
    ("table<float, C_order_> a;", "only column-major, owned, dense storage"),
    ("table<float, shared_> a;", "only column-major, owned, dense storage"),
    ("table<float, upper_triangular_> a;", "only column-major, owned, dense storage"),
    ("table<float, settings(_2D, column_major_, packed_)> a;", "only column-major, owned, dense storage"),
    ("table<float, settings(_2D, _3D)> a;", "one index base and one number of dimensions"),
    ("table<float, settings(C_index_, matlab_index_)> a;", "one index base and one number of dimensions"),
    ("table<float> a(of_size(4)); auto s = globalsum(a < 1.0f); (void)s;", "globalsum of a logical expression"),
    ("table<int32_t> a(of_size(4)); table<int32_t> b = pow(a, 2); (void)b;", "pow is defined for floating-point"),
]


@pytest.mark.parametrize("body", ACCEPTED)
def test_accepted_forms_compile(tmp_path, body):
    rc, err = compile_snippet(tmp_path, body)
    assert rc == 0, err[-3000:]


@pytest.mark.parametrize("body,message", REJECTED)
def test_rejected_forms_name_the_reason(tmp_path, body, message):
    rc, err = compile_snippet(tmp_path, body)
    assert rc != 0
    assert message in err, err[-3000:]
