"""CPU: the oracle's own tests run against an AddressSanitizer +
UndefinedBehaviorSanitizer build of the oracle (`make -C oracle asan-check`):
no out-of-bounds access, use after free or undefined behaviour in the host
restatement while it computes the pins, golden vectors and derivatives."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_address_and_ub_sanitizers():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ / make not available")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan-check"], capture_output=True, text=True,
                       timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "failed" not in r.stdout, tail
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, tail
