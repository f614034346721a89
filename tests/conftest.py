import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "trhepd-opt_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
REF = os.path.join(REPO, "tests", "golden", "ref")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def engine():
    from trhepd import lib
    e = lib.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="session")
def mixed_engine():
    from trhepd import lib
    e = lib.Engine(0, precision=lib.PREC_MIXED)
    yield e
    e.close()
