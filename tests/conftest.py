import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_single_thread():
    # deterministic summation order in the oracle's OpenMP reductions
    import oracle_ffi

    oracle_ffi.set_threads(1)
    yield
