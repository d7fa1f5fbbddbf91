import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer parity runs")


@pytest.fixture(scope="session")
def oracle_mod():
    import hfo
    hfo.build()
    return hfo
