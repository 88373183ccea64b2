import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


@pytest.fixture(scope="session")
def ref_lib():
    import oracle_lib

    lib = oracle_lib.ref()
    if lib is None:
        pytest.skip("flang-compiled reference routines (oracle/_ref) not built")
    return lib
