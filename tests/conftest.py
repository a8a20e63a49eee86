import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "finetoolsflexstructures.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C ABI")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the GPU tests (fails loudly without a device)."""
    from flexshell import _lib
    print(f"[gpu] creating context, lib={_lib.LIB_PATH}", flush=True)
    c = _lib.default_context()
    yield c
