import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libcourant_ref.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def cem():
    """The product engine (libcourant.so)."""
    from courant_amd import lib
    return lib.Courant()


@pytest.fixture(scope="session")
def ref():
    """The compiled reference core, as a CPU checker (built by oracle/Makefile)."""
    from courant_amd import lib
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference oracle/_ref not built")
    return lib.Courant(REF_LIB, extensions=False)


@pytest.fixture(scope="session")
def gpu_available():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        ok = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("GPU test selected but no HIP device is visible")
    return True
