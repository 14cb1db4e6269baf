import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "parties-vof-ibm_amd"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def gpu_available():
    try:
        import pvof
        import ctypes as C
        lib = pvof.load_lib()
        h = C.c_void_p()
        p = pvof.default_params(NXM=16, NYM=16)
        rc = lib.pv_create(C.byref(p), 0, C.byref(h))
        if rc == 0:
            lib.pv_destroy(h)
        return rc == 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("gpu test selected but no gfx950 device / library: the HIP path must run")
    return True
