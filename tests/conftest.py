import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import build, oracle
    if not os.path.exists(oracle.LIB):
        build.build()
    return oracle


@pytest.fixture(scope="session")
def af():
    """The product package with its HIP library loaded (GPU tests only)."""
    import aeroflux_amd
    aeroflux_amd.load_library()
    return aeroflux_amd
