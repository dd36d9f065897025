"""CPU: the legacy scene format (OldSceneFormat and the assembly of the
reference's legacy constructor, particlesolver.cuh:218-342, 395-710) as the
C++ shim restates it (brass-pan_amd/cpp/brasspan/oldscene.h).  The check is a
host-only C++ program (tests/cpp/legacy_assembly.cpp) compiled here."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_legacy_assembly(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists("/opt/rocm/include/hip/hip_vector_types.h"):
        pytest.skip("g++ or the HIP headers are not available")
    exe = tmp_path / "legacy_assembly"
    # the shim's headers need only HIP's vector types (as the Makefile's examples)
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I", os.path.join(ROOT, "brass-pan_amd", "cpp"),
                        os.path.join(ROOT, "tests", "cpp", "legacy_assembly.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "legacy assembly ok" in r.stdout
