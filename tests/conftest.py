import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def engine():
    """One qk_ctx on cuda device 0 for the whole GPU session.  A missing GPU or
    library is an error, not a skip: the HIP path must be the one that runs."""
    from quantkos_amd import Engine
    e = Engine(0)
    yield e
    e.close()
