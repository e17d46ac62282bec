import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wedge-parallel-triangle-counting_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        fx = json.load(f)
    for f_ in fx:
        f_["path"] = os.path.join(GOLDEN, f_["file"])
    return fx


@pytest.fixture(scope="session")
def gpu_ctx():
    import wtc
    if wtc.device_count() < 1:
        raise RuntimeError("GPU test collected but no HIP device is visible")
    ctx = wtc.Context(0)
    yield ctx
    ctx.close()
