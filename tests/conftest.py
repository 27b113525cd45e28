import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def capi():
    from tests.capi import CApi

    return CApi()


@pytest.fixture(scope="session")
def gpu(capi):
    """The C ABI, on a machine with a GPU (skips otherwise)."""
    if capi.device_count() < 1:
        pytest.skip("no GPU visible")
    return capi


@pytest.fixture(scope="session")
def golden():
    from tests.golden import load_all

    return load_all()


@pytest.fixture(scope="session")
def reference():
    from oracle import oracle as O

    R = O.load_reference()
    if R is None:
        pytest.skip("reference build (oracle/_ref) not present")
    return R
