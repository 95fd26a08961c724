"""Backend interface on the device: tests/test_backends.py:14-66, 131-144
and tests/test_ops.py:37-50 of the reference, for the HIP platform."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_setter(backend):
    backend.set_device("/GPU:0")
    assert backend.device == "/GPU:0"
    with pytest.raises(ValueError):
        backend.set_device("/CPU:0")
    with pytest.raises(ValueError):
        backend.set_device("/GPU:64")
    backend.set_device("/GPU:0")


def test_thread_setter(backend):
    original = backend.nthreads
    backend.set_threads(1)
    assert backend.nthreads == 1
    backend.set_threads(original)
    assert backend.nthreads == original


@pytest.mark.parametrize("array_type", [None, "float32", "float64", "complex64", "complex128"])
def test_cast(backend, array_type):
    target = np.random.default_rng(1).random(10)
    final = backend.to_numpy(backend.cast(target, dtype=array_type))
    np.testing.assert_allclose(final, target, rtol=1e-6 if array_type in ("float32", "complex64") else 1e-15)


def test_to_numpy(backend):
    x = [0, 1, 2]
    target = backend.to_numpy(backend.cast(x))
    final = backend.to_numpy(np.array(x))
    np.testing.assert_allclose(final, target)


def test_metabackend_list_available():
    from qibojit_amd.backends import PLATFORMS, MetaBackend

    available = MetaBackend().list_available()
    assert set(available) == set(PLATFORMS)
    assert available["hip"] is True
    assert MetaBackend.load("hip").platform == "hip"


@pytest.mark.parametrize("density_matrix", [False, True])
def test_plus_state(backend, dtype, density_matrix):
    backend.set_dtype(dtype)
    n = 4
    dims = 2**n
    final = backend.to_numpy(backend.plus_state(n, density_matrix=density_matrix))
    shape = (dims, dims) if density_matrix else (dims,)
    norm = dims if density_matrix else np.sqrt(dims)
    assert final.shape == shape and final.dtype == np.dtype(dtype)
    np.testing.assert_allclose(final, np.ones(shape) / norm, atol=1e-7 if dtype == "complex64" else 1e-15)
