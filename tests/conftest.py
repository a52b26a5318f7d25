import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); parity tests through the C ABI")


@pytest.fixture(scope="session")
def oracle():
    """The oracle's ctypes module (test infrastructure), built on demand."""
    from oracle import cpuref
    if not os.path.exists(cpuref.LIB_PATH):
        cpuref.build()
    return cpuref


@pytest.fixture(scope="session")
def ctx():
    from radiantium_amd.native import Context
    c = Context(0)
    yield c
    c.close()
