"""Register and scratch budget of the HIP kernels, read from the compiler's resource remarks for
gfx950 (no GPU needed): no kernel may spill to scratch memory, and the forward sweeps must keep
8 waves per SIMD (at most 64 VGPRs). A regression here shows up on the GPU as a many-fold
slowdown, not as a wrong result."""

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _remarks(src):
    from fdtdx_amd import build as B

    cmd = [B.hipcc(), *B.HIPCC_FLAGS, "--cuda-device-only", "-c", str(src), "-o", os.devnull,
           "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        for key in ("VGPRs", "ScratchSize [bytes/lane]", "Occupancy [waves/SIMD]"):
            m = re.search(re.escape(key) + r": (\d+)", line)
            if m and name:
                out[name][key] = int(m.group(1))
    return out


@pytest.fixture(scope="module")
def resources():
    from fdtdx_amd import build as B

    if not os.path.exists(B.hipcc()) and subprocess.run(["which", B.hipcc()], capture_output=True).returncode:
        pytest.skip("hipcc not available")
    srcs = [B.CSRC / s for s in B.SOURCES if (B.CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        parts = list(ex.map(_remarks, srcs))
    res = {}
    for p in parts:
        res.update(p)
    return res


def test_no_kernel_spills_to_scratch(resources):
    assert resources, "no kernel resource remarks parsed"
    bad = {k: v for k, v in resources.items() if v.get("ScratchSize [bytes/lane]", 0) != 0}
    assert not bad, bad


def test_forward_sweeps_keep_full_occupancy(resources):
    sweeps = {k: v for k, v in resources.items() if "k_update_E" in k or "k_update_H" in k}
    assert len(sweeps) >= 10
    for k, v in sweeps.items():
        assert v["VGPRs"] <= 64 and v["Occupancy [waves/SIMD]"] >= 8, (k, v)
