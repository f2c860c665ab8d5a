import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def gpu_available():
    """True when the engine library loads and sees a device."""
    try:
        from picca_amd import _native
        import ctypes
        lib = _native.load()
        count = ctypes.c_int32(0)
        return lib.picca_device_count(ctypes.byref(count)) == 0 and count.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def engine():
    """The process-wide engine on device 0 (GPU tests only; fails loudly without one)."""
    from picca_amd.engine import get_engine
    return get_engine(0)
