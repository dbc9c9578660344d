import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def gpu_available():
    """True when the HIP library loads and reports a device."""
    try:
        from guanaco_amd import _capi
        return _capi.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def capi():
    """The product's C-ABI library; on a GPU run it must load."""
    from guanaco_amd import _capi
    if _capi.device_count() <= 0:
        pytest.fail("no HIP device visible to the C-ABI library")
    return _capi
