import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "xrt-rs_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (runs on the GPU box)")


@pytest.fixture(scope="session")
def ctx():
    import xrtk
    c = xrtk.GpuContext.shared()
    if c is None:
        raise RuntimeError("a gpu-marked test ran without a gfx950 device")
    return c
