"""Workgroup barriers stay in the outermost loop of the persistent kernels
(CPU only: compiles the device code to gfx950 assembly and reads it).

A barrier that the optimizer moves into a nested loop can be re-entered by
some lanes of a wave while another lane waits at that loop's exit: the
workgroup then never leaves the loop.  In batch_persistent_k this happened
when the slow-tile branch's `threadIdx.x == 0` block was threaded into the
loop head's `threadIdx.x == 0` work fetch; a barrier between the two blocks
prevents it.  This test reads the loop depth the compiler annotates on each
block and fails if any barrier of the listed kernels sits deeper than their
persistent loop.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def barrier_depths(asm, kernel_prefix):
    out = {}
    for name in re.findall(r"^(_Z\S+):\s", asm, re.M):
        if kernel_prefix not in name:
            continue
        i = asm.index(name + ":")
        j = asm.find(".Lfunc_end", i)
        depth, depths = 0, []
        for line in asm[i:j].split("\n"):
            m = re.search(r"Depth=(\d+)", line)
            if line.startswith(".LBB") or line.startswith("; %bb"):
                depth = int(m.group(1)) if m else 0
            elif m and line.strip().startswith(";"):
                depth = int(m.group(1))
            if "s_barrier" in line:
                depths.append(depth)
        out[name] = depths
    return out


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not available")
def test_batch_barriers_in_outer_loop(tmp_path):
    sys_path = os.path.join(ROOT, "nt2_amd")
    import sys
    sys.path.insert(0, ROOT)
    from nt2_amd.build import HIP_FLAGS
    s_file = tmp_path / "batch.s"
    cmd = [HIPCC] + [f for f in HIP_FLAGS if f != "-fPIC"] + [
        "--cuda-device-only", "-S", os.path.join(sys_path, "csrc", "nt2h_batch.hip"), "-o", str(s_file)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    asm = s_file.read_text()
    for prefix in ("batch_persistent_k", "batch_slow_k"):
        found = barrier_depths(asm, prefix)
        assert found, prefix
        for name, depths in found.items():
            assert depths and max(depths) == 1, (name, depths)
