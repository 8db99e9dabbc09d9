import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built librpx.so")


@pytest.fixture(scope="session")
def example_rsb():
    with open(os.path.join(GOLDEN, "example.rsb"), "rb") as f:
        return f.read()


@pytest.fixture(scope="session")
def rpx():
    from repackshot_amd import native
    native.library()
    return native


@pytest.fixture(scope="session")
def gpu(rpx):
    if rpx.device_count() == 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return rpx
