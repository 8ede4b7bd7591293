import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    O.set_threads(min(8, os.cpu_count() or 1))
    return O


@pytest.fixture(scope="session")
def ctx64():
    from libprofit_amd.model import Context
    c = Context(0, "fp64")
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx32():
    from libprofit_amd.model import Context
    c = Context(0, "fp32")
    yield c
    c.close()
