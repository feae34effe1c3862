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
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built device library")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the whole GPU session (tests run in one process)."""
    from mxc_amd.device import Context, device_count

    if device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    c = Context(0)
    yield c
    c.close()
