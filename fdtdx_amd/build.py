"""Build the HIP engine library (libfdtdx_hip.so) in-tree for gfx950.

The library is compiled from fdtdx_amd/csrc/*.hip with hipcc. Floating-point contraction is off
so the kernels evaluate the reference's float32 expressions in the reference's order, and the
packed-FP32 target feature is disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libfdtdx_hip.so"
SOURCES = ["plan.hip", "fwd_kernels.hip", "det_kernels.hip", "adj_kernels.hip", "fused_kernels.hip",
           "march_kernels.hip", "calib_kernels.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if cand.startswith("/") and Path(cand).exists():
            return cand
    return "hipcc"


def build(verbose: bool = False, jobs: int = 8) -> Path:
    """Compile every source to an object (in parallel) and link the shared library."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    procs = []
    objs = []
    for src in srcs:
        obj = objdir / (src.stem + ".o")
        objs.append(obj)
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {src.name}")
        if verbose and out:
            sys.stderr.write(out.decode(errors="replace"))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr.decode(errors="replace"))
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


def build_plan_asan(verbose: bool = False) -> Path:
    """Build the host-side AddressSanitizer driver of the plan builder (tests/native/plan_asan.cpp):
    plan.hip and the driver are compiled with the host code under AddressSanitizer (the device code
    is not instrumented) and linked with the kernel objects of build(). Output:
    tests/native/plan_asan, run on a GPU by tests/test_gpu_plan_asan.py."""
    objdir = PKG / "_build"
    native = PKG.parent / "tests" / "native"
    asan = "-Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer -g".split()
    steps = [
        [hipcc(), *HIPCC_FLAGS, *asan, "-c", str(CSRC / "plan.hip"), "-o", str(objdir / "plan_asan_plan.o")],
        [hipcc(), *HIPCC_FLAGS, *asan, "-I", str(PKG.parent / "include"), "-c", str(native / "plan_asan.cpp"),
         "-o", str(objdir / "plan_asan_main.o")],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr.decode(errors="replace"))
            raise RuntimeError("hipcc failed on the AddressSanitizer driver")
    kernels = [objdir / (Path(s).stem + ".o") for s in SOURCES if s != "plan.hip"]
    exe = native / "plan_asan"
    cmd = [hipcc(), "--offload-arch=gfx950", *"-Xarch_host -fsanitize=address".split(),
           str(objdir / "plan_asan_main.o"), str(objdir / "plan_asan_plan.o"), *map(str, kernels), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr.decode(errors="replace"))
        raise RuntimeError("link of the AddressSanitizer driver failed")
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
