import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from tests import _oracle
    return _oracle


def gpu_available():
    try:
        from scaph_amd import _lib
        import ctypes
        c = ctypes.c_int(0)
        return _lib.hip().scaph_device_count(ctypes.byref(c)) == 0 and c.value > 0
    except Exception:
        return False
