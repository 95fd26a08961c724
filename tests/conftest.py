"""Fixtures: the HIP backend under test, the CPU port of NumbaBackend and
the dtype parametrisation of the reference's conftest (tests/conftest.py:35-39)."""

import os
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Tolerances of the reference's gate tests (tests/test_gates.py:15) and the
# parity bar of this backend against the CPU port (1e-12 for complex128,
# 1e-5 for complex64).
ATOL = {"complex64": 1e-4, "complex128": 1e-10}
PARITY = {"complex64": 1e-5, "complex128": 1e-12}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libqjhip.so")


@pytest.fixture(scope="session")
def hip_backend():
    """The HIP backend; raises (no skip) when the device or library is missing."""
    from qibojit_amd.backends.hip import HipBackend

    return HipBackend()


@pytest.fixture
def backend(hip_backend):
    hip_backend.set_dtype("complex128")
    hip_backend.set_threads(len(os.sched_getaffinity(0)))
    yield hip_backend
    hip_backend.set_dtype("complex128")


@pytest.fixture(scope="session")
def port_backend():
    from oracle import port

    return port.PortBackend(nthreads=min(8, len(os.sched_getaffinity(0))))


@pytest.fixture
def rng(request):
    # stable per-test seed derived from the test id
    return np.random.default_rng(zlib.crc32(request.node.nodeid.encode()))


def pytest_generate_tests(metafunc):
    if "dtype" in metafunc.fixturenames:
        metafunc.parametrize("dtype", ["complex128", "complex64"])
