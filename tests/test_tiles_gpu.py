"""Fused tile executor on the GPU: the kernel against the documented pass
semantics (numpy emulator), and fused circuit execution against the CPU
port of NumbaBackend and closed forms."""

import ctypes

import numpy as np
import pytest

from oracle import gate_matrices as gm
from oracle import port
from qibojit_amd import _native, tiles
from qibojit_amd._qibo import Circuit, gates
from tests.conftest import PARITY
from tests.test_tiles import _random_circuit
from tests.tile_emulator import apply_tile_pass

pytestmark = pytest.mark.gpu


def _qft(n):
    c = Circuit(n)
    for i in range(n):
        c.add(gates.H(i))
        for j in range(i + 1, n):
            c.add(gates.CU1(j, i, np.pi / 2 ** (j - i)))
    for i in range(n // 2):
        c.add(gates.SWAP(i, n - i - 1))
    return c


@pytest.mark.parametrize("n,low,nhigh", [(10, 5, 5), (12, 6, 6), (11, 1, 7), (13, 4, 8)])
def test_tile_kernel_matches_emulator(backend, n, low, nhigh, dtype, rng):
    backend.set_dtype(dtype)
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    if (2 ** (low + nhigh)) * np.dtype(npdt).itemsize > 65536:
        pytest.skip("tile over 64 KiB")
    circuit = _random_circuit(n, 2, rng)
    ops = [tiles.gate_to_op(port.custom_matrix, g, n) for g in circuit]
    ops = [op for op in ops if tiles.fits_alone(op, low, nhigh)]
    passes = tiles.plan(ops, n, low, nhigh)
    psi = gm.random_state(n, rng, npdt)
    state = backend.cast(psi, dtype=dtype, copy=True)
    want = psi.copy()
    for p in passes:
        enc, data, high = tiles.encode(p, n, npdt)
        want = apply_tile_pass(want, n, p.low, high, enc, data, npdt)
        harr = (ctypes.c_int32 * max(1, len(high)))(*high)
        dbuf = ctypes.create_string_buffer(data, max(1, len(data)))
        _native.call("qj_apply_tile_pass", ctypes.c_void_p(state.data_ptr()), _native.QJ_COMPLEX64 if dtype == "complex64" else _native.QJ_COMPLEX128,
                     n, p.low, ctypes.cast(harr, ctypes.POINTER(ctypes.c_int32)), len(high),
                     enc.ctypes.data_as(ctypes.c_void_p), len(enc), ctypes.cast(dbuf, ctypes.c_void_p), len(data), backend._stream())
    got = backend.to_numpy(state)
    np.testing.assert_allclose(got, want, atol=PARITY[dtype] * 10, rtol=0)


def test_tile_pass_without_ops_is_identity(backend, rng):
    """Regression: an empty op list once made the staging copy read from a
    null host pointer; the pass must just copy every tile back in place."""
    psi = gm.random_state(12, rng)
    state = backend.cast(psi, copy=True)
    p = tiles.Pass(5)
    p.high = [6, 8, 11]
    tiles.run_passes(_native.call, ctypes.c_void_p(state.data_ptr()), _native.QJ_COMPLEX128, 12, [p],
                     np.complex128, backend._stream())
    np.testing.assert_array_equal(backend.to_numpy(state), psi)


def test_tile_pass_rejects_bad_ops(backend):
    state = backend.zero_state(8)
    op = np.zeros(1, dtype=tiles.TILE_OP_DTYPE)
    op["kind"], op["k"], op["target"][0][0] = tiles.TILE_DENSE, 1, 9  # target outside the tile
    data = np.eye(2, dtype=np.complex128).tobytes()
    harr = (ctypes.c_int32 * 2)(5, 6)
    with pytest.raises(ValueError):
        _native.call("qj_apply_tile_pass", ctypes.c_void_p(state.data_ptr()), _native.QJ_COMPLEX128, 8, 3,
                     ctypes.cast(harr, ctypes.POINTER(ctypes.c_int32)), 2, op.ctypes.data_as(ctypes.c_void_p), 1,
                     data, len(data), backend._stream())


@pytest.mark.parametrize("n", [3, 10, 16])
def test_fused_random_circuit(backend, port_backend, n, dtype, rng):
    backend.set_dtype(dtype)
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    c = Circuit(n)
    c.add(_random_circuit(n, 3, rng))
    psi = gm.random_state(n, rng, npdt)
    want = port_backend.execute_circuit(c, psi)
    got = backend.to_numpy(backend.execute_circuit(c, psi.copy()).state())
    np.testing.assert_allclose(got, want, atol=PARITY[dtype] * 10, rtol=0)


def test_fused_density_matrix_circuit(backend, port_backend, rng):
    n = 5
    c = Circuit(n, density_matrix=True)
    c.add(_random_circuit(n, 2, rng))
    rho = gm.random_density_matrix(n, rng)
    want = port_backend.execute_circuit(c, rho)
    got = backend.to_numpy(backend.execute_circuit(c, rho.copy()).state())
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)


@pytest.mark.parametrize("n", [8, 15, 20])
def test_fused_qft_matches_dft(backend, n, dtype, rng):
    backend.set_dtype(dtype)
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    psi = gm.random_state(n, rng, npdt)
    got = backend.to_numpy(backend.execute_circuit(_qft(n), psi.copy()).state())
    want = np.fft.ifft(psi.astype(np.complex128)) * np.sqrt(2**n)
    np.testing.assert_allclose(got, want, atol=PARITY[dtype] * 10 if dtype == "complex64" else 1e-12, rtol=0)


def test_fused_qft_large_uniform_superposition(backend):
    """QFT|0...0> is the uniform superposition (26 qubits, complex64)."""
    backend.set_dtype("complex64")
    n = 26
    out = backend.execute_circuit(_qft(n)).state()
    amp = 2 ** (-n / 2)
    err = float((out - amp).abs().max())
    assert err < 1e-5 * amp * 2**3, err
    assert abs(backend.norm_squared(out) - 1) < 1e-5


def test_fusion_switch_gives_same_state(backend, rng):
    n = 12
    c = Circuit(n)
    c.add(_random_circuit(n, 2, rng))
    psi = gm.random_state(n, rng)
    backend.fusion = False
    try:
        a = backend.to_numpy(backend.execute_circuit(c, psi.copy()).state())
    finally:
        backend.fusion = True
    b = backend.to_numpy(backend.execute_circuit(c, psi.copy()).state())
    np.testing.assert_allclose(a, b, atol=1e-12, rtol=0)
