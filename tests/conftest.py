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
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gb():
    import libgpublas_amd as gb
    gb.library()
    return gb


@pytest.fixture(scope="session")
def ora():
    import _oracle
    return _oracle.load()


@pytest.fixture()
def gpu(gb):
    """GPU path forced (no host fall-through); checks the device is there."""
    if not gb.device_ready():
        pytest.fail("no gfx950 device visible to libgpublas")
    old = gb.get_cutoff()
    gb.set_cutoff(0)
    yield gb
    gb.set_cutoff(old)
