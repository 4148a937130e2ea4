import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the gfx950 path")
    config.addinivalue_line("markers", "ref: needs the reference build oracle/_ref (built where the reference checkout exists)")


@pytest.fixture(scope="session")
def lib():
    from lighter_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def orc():
    from oracle.checkers import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle.checkers import Reference, reference_available
    if not reference_available():
        pytest.skip("reference build oracle/_ref/liblighter_ref.so not present")
    return Reference()


@pytest.fixture(scope="session")
def libc():
    import ctypes
    return ctypes.CDLL(None)
