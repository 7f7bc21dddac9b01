import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path of libksh.so)")


@pytest.fixture(scope="session")
def ksh():
    """The GPU library; GPU tests fail loudly when it is missing."""
    from gsstructuralanalysis_amd import _lib
    return _lib.load()
