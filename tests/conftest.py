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
    from eating_one_amd import _abi
    from eating_one_amd import build as eob

    # the GPU suite must exercise the library built from these sources
    built = _abi.load().eo_build_source_hash().decode()
    assert built == eob.source_hash(), f"libeojoin.so ({built}) is stale against the tree ({eob.source_hash()})"
    c = eo.Context(0)
    yield c
    c.close()
