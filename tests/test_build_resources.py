"""Kernel resource checks of the HIP sources, on the CPU (hipcc cross-compiles
for gfx950 without a GPU).  A kernel that spills to scratch memory runs much
slower than it should: a GEMM tile whose K-slab coordinates were assigned in
branches once kept them in scratch and doubled every FP64 GEMM's time.  The
compiler's resource remarks name each kernel's scratch bytes per lane; every
kernel of the matrix-core GEMMs, the Jacobi SVD and the core kernels must have
none."""
import os
import re
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "tensor_train_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _scratch_by_kernel(src):
    out = subprocess.run(
        [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-c", "-o", os.devnull, os.path.join(CSRC, src), "-Rpass-analysis=kernel-resource-usage"],
        capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    res, name = {}, None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            res[name] = int(m.group(1))
    return res


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src", ["tta_kernels.hip", "tta_jacobi.hip"])
def test_no_scratch_memory(src):
    res = _scratch_by_kernel(src)
    assert len(res) >= 10  # the remarks were parsed
    spilled = {k: v for k, v in res.items() if v > 0}
    assert not spilled, spilled
