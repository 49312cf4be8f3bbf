import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_built():
    from aquarius_amd._abi import LIB_PATH
    from oracle import binding
    if not binding.LIB_PATH.exists():
        binding.build()
    if not LIB_PATH.exists():
        from aquarius_amd import build
        build.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_available():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        rc = hip.hipGetDeviceCount(ctypes.byref(n))
        return rc == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(autouse=True)
def _require_gpu(request, gpu_available):
    if request.node.get_closest_marker("gpu") and not gpu_available:
        pytest.fail("GPU test collected on a machine without a GPU (run with -m 'not gpu')")
