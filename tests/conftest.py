import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU oracle cases")


@pytest.fixture(scope="session")
def gpu():
    import ctypes
    from hs_sfm_amd import abi
    lib = abi.load_library()  # raises when libhsba.so is missing: no fallback
    n = ctypes.c_int(0)
    lib.hsba_device_count(ctypes.byref(n))
    if n.value == 0:
        pytest.skip("no GPU visible")
    return lib
