import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libsidproj.so")


@pytest.fixture(scope="session")
def oracle_lib():
    """The C oracle (built by oracle/build_oracle.py); built on demand here."""
    import ctypes
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        from oracle.build_oracle import build
        build()
    return ctypes.CDLL(path)
