"""Register budget of the hot FP32 kernels, read from the compiler's
resource-usage remarks (hipcc -Rpass-analysis=kernel-resource-usage for
gfx950): the spill-free builds the launch tuner relies on must stay
spill-free and must not place arrays in scratch memory (a lane-dependent
index into a private array does that, and it multiplies the kernel's
memory traffic)."""
import os
import re
import shutil
import subprocess

import pytest

from qsource_amd import build

CSRC = os.path.join(os.path.dirname(build.LIB), "csrc")


def _resources(src):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    flags = [f for f in build.HIP_FLAGS if f != "-fPIC"]
    r = subprocess.run([hipcc] + flags + ["--cuda-device-only", "-Rpass-analysis=kernel-resource-usage",
                                          "-c", os.path.join(CSRC, src), "-o", os.devnull],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    out, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|VGPRs Spill|ScratchSize \[bytes/lane\]): (\d+)", line)
        if m and cur:
            out[cur][m.group(1)] = int(m.group(2))
    return out


def test_fp32_water_forces_spill_free():
    res = _resources("kernels_water.hip")
    # k_water_forces<float, MAPPED, PF, W=4, AA=0>
    names = [k for k in res if re.search(r"k_water_forcesIfLi[01]ELi[01]ELi4ELi0E", k)]
    assert len(names) == 4, names
    for k in names:
        assert res[k].get("VGPRs Spill", 0) == 0, (k, res[k])
        assert res[k].get("ScratchSize [bytes/lane]", 0) == 0, (k, res[k])
    # the list kernels in both precisions
    for k in [k for k in res if "k_water_lists" in k]:
        assert res[k].get("ScratchSize [bytes/lane]", 0) == 0, (k, res[k])


def test_solute_kernels_no_scratch():
    res = _resources("kernels_solute.hip")
    hot = [k for k in res if re.search(r"k_solute_forcesIf|k_solute_listsIf|k_q_forcesIf", k)]
    assert hot
    for k in hot:
        assert res[k].get("ScratchSize [bytes/lane]", 0) == 0, (k, res[k])
