import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


_gpu = None


def gpu_available():
    """True when the HIP runtime reports a device (counting devices does not create a context)."""
    global _gpu
    if _gpu is None:
        try:
            hip = ctypes.CDLL("libamdhip64.so")
            n = ctypes.c_int(0)
            _gpu = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
        except OSError:
            _gpu = False
    return _gpu


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    return True
