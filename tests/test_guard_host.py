"""The non-finite guard's logic on the host (bp_guard.h, shared with the
device scan k_first_nonfinite): a C++ program compiled here with g++, no HIP
and no device.  The device path itself is covered by
test_gpu_parity.py::test_nan_guard_names_the_particle."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_guard_logic_on_host(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path / "guard_check"
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "brass-pan_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "guard_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "guard ok" in r.stdout, r.stdout + r.stderr
