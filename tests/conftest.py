import importlib
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
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


@pytest.fixture(scope="session")
def fv():
    return importlib.import_module("fvdbm-jax_amd")


@pytest.fixture(scope="session")
def gpu(fv):
    """Fail (not skip) when a GPU test runs without a device or library."""
    from importlib import import_module
    lib = import_module("fvdbm-jax_amd._lib")
    n = lib.device_count()
    if n < 1:
        pytest.fail("GPU test selected but no HIP device is visible")
    return n
