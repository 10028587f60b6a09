"""pytest configuration: the `gpu` marker (tests that need an AMD GPU and the built HIP library) and the repository
root on sys.path, so that tests import the package and the oracle from the working tree."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def have_gpu():
    return os.path.exists('/dev/kfd')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU (MI355X) and the built HIP library')
    config.addinivalue_line('markers', 'slow: long CPU test')


def pytest_collection_modifyitems(config, items):
    if have_gpu():
        return
    skip = pytest.mark.skip(reason='no AMD GPU here (/dev/kfd absent)')
    for item in items:
        if 'gpu' in item.keywords:
            item.add_marker(skip)
