import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libeojoin.so on the device)")


@pytest.fixture(scope="session")
def ctx():
    """One library context on GPU 0 for the whole GPU session."""
    import eating_one_amd as eo

    c = eo.Context(0)
    yield c
    c.close()
