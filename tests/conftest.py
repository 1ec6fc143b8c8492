import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running case")


def gpu_available() -> bool:
    try:
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def lib():
    import sysu_cfd_parallel_amd as s

    return s.library()


@pytest.fixture(scope="session")
def olib():
    from oracle import pyoracle

    return pyoracle.library()


@pytest.fixture(autouse=True)
def _print_case(request):
    # name the case before any GPU step, so a box that dies still leaves the name behind
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
