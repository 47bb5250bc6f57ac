import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an AMD GPU (runs the HIP kernels)')


@pytest.fixture(scope='session')
def lib():
    """The device library; GPU tests fail loudly if it is missing."""
    from myokit_amd import _native
    return _native.lib()
