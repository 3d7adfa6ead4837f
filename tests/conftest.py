import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "badchimp-cpp_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def gpu_available():
    import lbamd
    try:
        ctx = lbamd.Context(lbamd.D2Q9, lbamd.FP64, 0)
    except lbamd.LBError as e:
        pytest.fail("GPU test on a machine without a usable HIP device: %s" % e)
    ctx.close()
    return True
