import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def mivf():
    """The product library; GPU tests fail loudly if it is missing."""
    import mestrado_amd as M
    M.lib()
    assert M.device_count() >= 1, "no HIP device visible"
    return M


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle as O
    O.lib()
    return O
