import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libb40c_hip.so on the device)")
    config.addinivalue_line("markers", "slow: full-size configuration (seconds of host work)")


@pytest.fixture(scope="session")
def lib():
    from back40computing_amd import _native
    return _native.load()


@pytest.fixture(scope="session")
def gpu():
    """Asserts a device is present: GPU tests never fall back to the host."""
    import back40computing_amd as b40c
    n = b40c.device_count()
    assert n >= 1, "GPU test requires a HIP device"
    return 0
