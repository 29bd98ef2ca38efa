import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU-only case")
    # PyTorch's ROCm wheel carries its own HIP runtime and libpll.so binds to
    # whichever libamdhip64.so.7 the process loaded first; torch only finds the
    # GPU through its own.  GPU tests use torch for device buffers, so it is
    # imported before any test loads libpll.so.
    if "not gpu" not in (config.getoption("markexpr") or ""):
        try:
            import torch  # noqa: F401
        except ImportError:
            pass


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pll_oracle
    pll_oracle.build()
    return pll_oracle
