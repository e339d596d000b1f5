import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (HIP device); parity vs the oracle")


def gpu_available():
    try:
        from oskar_amd import lib
        return lib.oskar_device_count_hip() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    """The loaded C-ABI library; the test fails (not skips) if it is missing
    or no device is visible, so a GPU run never passes on a fallback."""
    from oskar_amd import lib
    n = lib.oskar_device_count_hip()
    assert n > 0, "no HIP device visible"
    return lib
