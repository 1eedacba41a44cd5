import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device and the built libdwg_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU-oracle runs")


@pytest.fixture(scope="session")
def ctx():
    from dwg_amd.gpu import Context
    c = Context(0)
    yield c
    c.close()
