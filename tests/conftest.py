import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def gpu_lib():
    """libgemm_amd.so with a usable GPU; a GPU test fails (never falls back) if either is missing."""
    from gemm_avx512f_amd import _lib
    L = _lib.lib()
    rc = L.gemm_amd_init()
    assert rc == 0, "HIP init failed: " + _lib.last_error()
    return L
