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
    """One-cell sweeps: 8 waves per SIMD. Quad sweeps (four cells per lane, k_update_E4/H4): at
    least 5, i.e. 20 cells per lane slot in flight against the one-cell sweeps' 8."""
    quad = {k: v for k, v in resources.items() if "k_update_E4" in k or "k_update_H4" in k}
    sweeps = {k: v for k, v in resources.items() if ("k_update_E" in k or "k_update_H" in k) and k not in quad}
    assert len(sweeps) >= 10 and len(quad) >= 10
    for k, v in sweeps.items():
        assert v["VGPRs"] <= 64 and v["Occupancy [waves/SIMD]"] >= 8, (k, v)
    for k, v in quad.items():
        assert v["VGPRs"] <= 96 and v["Occupancy [waves/SIMD]"] >= 5, (k, v)


@pytest.mark.parametrize("ny,nz,zlo,zhi,zg", [(400, 200, 12, 188, 32), (45, 70, 5, 65, 7), (30, 24, 0, 19, 32),
                                              (17, 33, 6, 33, 4), (20, 16, 0, 16, 20)])
def test_forward_plane_order_is_a_permutation(ny, nz, zlo, zhi, zg, tmp_path):
    """The forward sweeps' plane order visits every cell of a plane exactly once, group by group
    of zg rows, the z-interior cells of a group before its z-slab cells."""
    import numpy as np

    from fdtdx_amd import build as B

    exe = str(tmp_path / "order")
    r = subprocess.run([B.hipcc(), "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-o", exe,
                        os.path.join(ROOT, "tests", "plane_order_harness.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe, str(ny), str(nz), str(zlo), str(zhi), str(zg)], capture_output=True, text=True)
    jk = np.array([list(map(int, l.split())) for l in out.stdout.split("\n") if l], dtype=np.int64)
    assert jk.shape == (ny * nz, 2)
    flat = jk[:, 0] * nz + jk[:, 1]
    assert np.array_equal(np.sort(flat), np.arange(ny * nz))
    edge = (jk[:, 1] < zlo) | (jk[:, 1] >= zhi)
    group = jk[:, 0] // zg
    assert np.all(np.diff(group) >= 0)
    for g in np.unique(group):
        e = edge[group == g]
        assert not np.any(np.diff(e.astype(int)) < 0)  # interior first, then the slab cells
