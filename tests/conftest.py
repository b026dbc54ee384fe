import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcaffe_hip.so)")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def oracle_built():
    from oracle import pyoracle
    pyoracle.lib()
    return True
