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
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libqdx.so")
    config.addinivalue_line("markers", "slow: long-running CPU reference runs")


@pytest.fixture(scope="session")
def ctx():
    import qdng_amd as Q
    c = Q.Context(0)
    yield c
    c.close()
