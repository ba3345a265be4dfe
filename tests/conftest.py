import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pagani-toolkit_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpagani kernels)")


@pytest.fixture(scope="session")
def ctx():
    import pagani_amd
    c = pagani_amd.Context(0)
    yield c
    c.close()


_HIP = None


def hip_d2h(ptr: int, nbytes: int) -> bytes:
    """Copy device memory to host through the HIP runtime (test helper)."""
    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipDeviceSynchronize.argtypes = []
    buf = ctypes.create_string_buffer(max(nbytes, 1))
    assert _HIP.hipDeviceSynchronize() == 0
    assert _HIP.hipMemcpy(buf, ctypes.c_void_p(ptr), nbytes, 2) == 0  # DeviceToHost
    return buf.raw[:nbytes]


def say(*a):
    """Name the case on stdout before a GPU step (flushed)."""
    print(*a, flush=True)
