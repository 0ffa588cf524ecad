import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X) and the built HIP library')
    config.addinivalue_line('markers', 'slow: long-running test')


@pytest.fixture(autouse=True)
def _reset_configuration():
    """Hermetic configuration per test (as the reference's tests/conftest.py:14-18)."""
    from hcipy_amd.config import Configuration
    Configuration().reset()
    yield
