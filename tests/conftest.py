import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def gpu_available():
    # through the library itself: importing torch here would load torch's
    # bundled HIP runtime first, and the library would then run on it
    try:
        from nanoshaper_amd import _lib
        return _lib.load().nsgpu_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def device_ok():
    if not gpu_available():
        pytest.skip("no HIP device")
    return True
