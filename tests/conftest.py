import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "geophyinv.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libgpi.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One libgpi context on device 0 for the whole GPU session."""
    from geophyinv import _lib
    ctx = _lib.Context(0)
    yield ctx
    ctx.close()
