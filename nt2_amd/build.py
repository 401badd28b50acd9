"""Builds the native pieces in-tree: libnt2h.so (hipcc, gfx950) from
nt2_amd/csrc, and the oracle (oracle/liboracle.so, g++) used by the tests.

    python -m nt2_amd.build            # everything
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 (plain float code only);
# -ffp-contract=off: NT2's arithmetic has no implicit contraction (explicit
# fused multiply-adds come only from the fast policy's __builtin_fma).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(HERE, "include"), "-I" + os.path.join(ROOT, "include")]

SOURCES = ["nt2h_core.hip", "nt2h_batch.hip", "nt2h_server.hip", "nt2h_dimwise.hip", "nt2h_demos.hip",
           "nt2h_compose.hip"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return p


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s + ".o") for s in srcs]

    def compile_one(i):
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, srcs[i]), "-o", objs[i]])

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(compile_one, range(len(srcs))))
    out = os.path.join(HERE, "libnt2h.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_frontend_tests():
    """Compile the C++ front-end programs under tests/cpp (user code written
    against nt2/table.hpp) with hipcc into tests/cpp/bin."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    if not os.path.isdir(src_dir):
        return []
    bin_dir = os.path.join(src_dir, "bin")
    os.makedirs(bin_dir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(src_dir) if f.endswith(".cpp"))
    outs = []

    def compile_one(f):
        exe = os.path.join(bin_dir, f[:-4])
        _run([HIPCC, "-x", "hip"] + HIP_FLAGS + [os.path.join(src_dir, f), "-o", exe,
              "-L" + HERE, "-lnt2h", "-Wl,-rpath," + HERE])
        return exe

    with ThreadPoolExecutor(max_workers=max(1, min(8, len(srcs)))) as ex:
        outs = list(ex.map(compile_one, srcs))
    return outs


def build_all():
    lib = build_lib()
    orc = build_oracle()
    fe = build_frontend_tests()
    return [lib, orc] + fe


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
