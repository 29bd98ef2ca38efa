import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU-only case")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pll_oracle
    pll_oracle.build()
    return pll_oracle
