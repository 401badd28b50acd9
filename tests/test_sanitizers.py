"""Host-side sanitizer runs of the serving scheduler (opt-in: set
NT2H_SANITIZERS=1; each build takes about a minute).  The library's sources
are compiled with ThreadSanitizer, then AddressSanitizer (LeakSanitizer on),
on the host pass only (-Xarch_host: no GPU code is instrumented), linked with
tests/cpp/server_stress.cpp and run on the host back end: no report, every
request ends and the budget returns to zero."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["nt2h_core", "nt2h_server", "nt2h_batch", "nt2h_dimwise", "nt2h_demos"]

pytestmark = pytest.mark.skipif(os.environ.get("NT2H_SANITIZERS") != "1",
                                reason="opt-in: NT2H_SANITIZERS=1 (builds the library twice)")


def _build(san, out):
    from concurrent.futures import ThreadPoolExecutor
    flags = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "nt2_amd", "include"), "-I" + os.path.join(ROOT, "include")]

    def one(f):
        subprocess.run([HIPCC] + flags + ["-Xarch_host", "-fsanitize=" + san, "-c",
                        os.path.join(ROOT, "nt2_amd", "csrc", f + ".hip"), "-o", os.path.join(out, f + ".o")],
                       check=True, capture_output=True)
    with ThreadPoolExecutor(max_workers=5) as ex:
        list(ex.map(one, SOURCES))
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-fsanitize=" + san, "-c",
                    os.path.join(HERE, "cpp", "server_stress.cpp"), "-o", os.path.join(out, "driver.o")],
                   check=True, capture_output=True)
    exe = os.path.join(out, "server_stress_" + san)
    subprocess.run([HIPCC, "-Xarch_host", "-fsanitize=" + san, os.path.join(out, "driver.o")] +
                   [os.path.join(out, f + ".o") for f in SOURCES] + ["-o", exe, "-lpthread"],
                   check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("san", ["thread", "address"])
def test_server_stress_sanitized(san):
    with tempfile.TemporaryDirectory() as d:
        exe = _build(san, d)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", TSAN_OPTIONS="halt_on_error=1")
        p = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=env)
        log = p.stdout + p.stderr
        assert "ThreadSanitizer" not in log and "AddressSanitizer" not in log and "LeakSanitizer" not in log, log[-4000:]
        assert p.returncode == 0, log[-4000:]
