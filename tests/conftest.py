import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (run on the MI355X box)")


def gpu_available():
    try:
        import cuopo3d_amd
        return cuopo3d_amd.load().opo_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    import cuopo3d_amd
    n = cuopo3d_amd.load().opo_device_count()
    if n < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return n
