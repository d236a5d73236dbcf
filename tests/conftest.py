import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_ffi
    oracle_ffi.lib()
    return oracle_ffi


@pytest.fixture(scope="session")
def engine():
    """The HIP engine on device 0.  Fails (does not skip) when the library or
    the device is missing: GPU tests must run the native path."""
    from geliosphere_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()
