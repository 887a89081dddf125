"""Host-side AddressSanitizer run of the plan builder (tests/native/plan_asan.cpp): a valid plan
with CPML, a dipole, flux / phasor / energy detectors and the interface record goes through a
forward and the reverse adjoint, and some fifty malformed descriptors are each rejected with a
status and a message. plan.hip's host code and the driver are built with AddressSanitizer (the
device code is not instrumented) by fdtdx_amd/build.py's build_plan_asan; leaks are checked
with the HIP and HSA runtimes' own allocations suppressed."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNTIMES = ["libamdhip64", "libhsa-runtime64", "libhsakmt", "libdrm", "libamd_comgr", "librocprofiler"]


def test_plan_builder_under_address_sanitizer(tmp_path):
    exe = os.path.join(ROOT, "tests", "native", "plan_asan")
    if not os.path.exists(exe):
        pytest.fail("tests/native/plan_asan is not built: run __graft_entry__.build()")
    supp = tmp_path / "lsan.supp"
    supp.write_text("".join(f"leak:{lib}\n" for lib in RUNTIMES))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               LSAN_OPTIONS=f"suppressions={supp}:print_suppressions=0")
    print("case plan_asan", exe, flush=True)
    r = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=180)
    print(r.stdout[-4000:], flush=True)
    print(r.stderr[-8000:], flush=True)
    assert r.returncode == 0, f"plan_asan exited with {r.returncode}"
    assert "cases passed" in r.stdout.splitlines()[-1]
