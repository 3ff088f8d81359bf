"""Every gfx950 kernel of the library runs without scratch (private-segment)
memory: the compiler's resource report for each device source shows
ScratchSize 0 for every kernel.  A kernel that spills needs the runtime to
allocate scratch backing at dispatch time, which fails when the device
memory is held by resident DMRG states, and the queue then dies with
HSA_STATUS_ERROR_OUT_OF_RESOURCES."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "symmetry-adapted-dmrg-library_amd")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.parametrize("src", ["kernels.hip", "eigen.hip"])
def test_no_kernel_uses_scratch(src, tmp_path):
    if not os.path.exists(HIPCC) or shutil.which("ld.lld", path="/opt/rocm/llvm/bin") is None:
        pytest.skip("hipcc not available")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-Rpass-analysis=kernel-resource-usage", "-c",
           os.path.join(PKG, "csrc", src), "-o", str(tmp_path / "dev.o")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    names = re.findall(r"Function Name: (\S+)", r.stderr)
    scratch = [int(x) for x in re.findall(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)]
    assert names and len(names) == len(scratch)
    spilled = [(n, s) for n, s in zip(names, scratch) if s > 0]
    assert not spilled, spilled
