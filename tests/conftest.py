import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) GPU; runs the HIP kernels')


@pytest.fixture(scope='session')
def oracle():
    from tests._cpu_libs import oracle as _o
    return _o()


@pytest.fixture(scope='session')
def reference():
    from tests._cpu_libs import reference as _r
    r = _r()
    if r is None:
        pytest.skip('compiled reference (oracle/_ref) not built here')
    return r


@pytest.fixture(scope='session')
def reference_zi():
    from tests._cpu_libs import reference_zi as _r
    r = _r()
    if r is None:
        pytest.skip('zero-initialising reference build (oracle/_ref/libqtxref_zi.so) not built here')
    return r


@pytest.fixture(scope='session')
def qtx():
    """The product's C ABI.  On a GPU run a missing library is an error, not a skip."""
    from tests._qtx import Qtx
    q = Qtx()
    yield q
    q.close()
