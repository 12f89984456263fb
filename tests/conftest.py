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
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")
    config.addinivalue_line("markers", "slow: long CPU test (oracle convergence runs)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (test infrastructure), built from oracle/ if needed."""
    from oracle import pyoracle

    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device; GPU tests fail (not skip) when it is missing."""
    from ppkmhd_amd._lib import lib

    L = lib()
    n = L.ppk_mhd_device_count()
    assert n > 0, "no HIP device visible to a test marked gpu"
    return L
