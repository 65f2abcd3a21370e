"""Test configuration: the `gpu` marker and shared fixtures.

GPU tests (`-m gpu`) call the HIP library through its C ABI and compare
with the CPU oracle; CPU tests (`-m "not gpu"`) pin the oracle and check
host-side logic and the library's exported interface.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def capi():
    from cuphu_amd import _capi
    return _capi


@pytest.fixture(scope="session")
def gpu_ready(capi):
    n = capi.lib.cuphu_gpu_count()
    if n < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return n


@pytest.fixture()
def rng():
    return np.random.default_rng(1234)
