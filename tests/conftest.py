import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), '..'))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an MI355X (runs the HIP kernels)')
    config.addinivalue_line(
        'markers', 'slow: long-running')


@pytest.fixture(scope='session')
def golden():
    """golden('fd') -> dict-like view of tests/golden/fd_golden.npz."""
    cache = {}

    def load(name):
        if name not in cache:
            path = os.path.join(ROOT, 'tests', 'golden', f'{name}_golden.npz')
            cache[name] = np.load(path, allow_pickle=False)
        return cache[name]
    return load
