import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: larger problem sizes")


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when the device is missing: a GPU test must run the HIP path."""
    from compressiontests_amd import _lib
    n = _lib.device_count()
    assert n > 0, "no GPU visible to libmgx.so"
    return n
