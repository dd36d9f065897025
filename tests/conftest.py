import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("brass-pan_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def have_oracle():
    import oracle_py
    if not os.path.exists(oracle_py.LIB_PATH):
        pytest.skip("oracle/lib/liboracle.so not built (run __graft_entry__.build())")
    return True
