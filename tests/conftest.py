import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


def gpu_available():
    try:
        from h2_amd import _lib
        import ctypes
        lib = _lib.load()
        ctx = ctypes.c_void_p()
        rc = lib.h3d_create(ctypes.byref(ctx), 3, 3, 3, 0)
        if rc == 0:
            lib.h3d_destroy(ctx)
            return True
    except Exception:
        pass
    return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.load()
    return O
