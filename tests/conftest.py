import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')
    config.addinivalue_line('markers', 'slow: long-running CPU test')


@pytest.fixture(scope='session')
def oracle():
    from oracle.cpu import OracleKernels
    return OracleKernels()


@pytest.fixture(scope='session')
def ref():
    from oracle.cpu import RefKernels
    if not RefKernels.available():
        pytest.skip('oracle/_ref (the reference C build) is not available')
    return RefKernels()


@pytest.fixture(scope='session')
def gpu():
    """The GPU helper; fails (does not skip) when no device is visible, so that a
    GPU run without a device is reported as an error rather than silently passing."""
    from gpu_helpers import GPU
    return GPU()
