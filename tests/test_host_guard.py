"""No C++ exception crosses the C ABI: the library's entry points run their
bodies under host_guard and its host builders join their thread pools on every
way out (elfel.jl_amd/csrc/common.h).  tests/host_guard_check.cpp exercises
both in a host-only program (no device calls: runs without a GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not installed")
def test_exceptions_become_error_codes(tmp_path):
    hipcc = HIPCC if os.path.exists(HIPCC) else shutil.which("hipcc")
    exe = tmp_path / "host_guard_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "elfel.jl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host_guard_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
