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
from tests.tile_emulator import apply_bit_swaps, apply_tile_pass

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


def _code(dtype):
    return _native.QJ_COMPLEX64 if dtype == "complex64" else _native.QJ_COMPLEX128


def _run_and_emulate(backend, passes, n, dtype, psi):
    """Runs the passes through the C-ABI and the numpy emulator."""
    npdt = psi.dtype
    state = backend.cast(psi, dtype=dtype, copy=True)
    want = psi.copy()
    for p in passes:
        if isinstance(p, tiles.SwapPass):
            want = apply_bit_swaps(want, n, p.pairs)
            flat = (ctypes.c_int32 * (2 * len(p.pairs)))(*[b for pr in p.pairs for b in pr])
            _native.call("qj_apply_bit_swaps", ctypes.c_void_p(state.data_ptr()), _code(dtype), n,
                         ctypes.cast(flat, ctypes.POINTER(ctypes.c_int32)), len(p.pairs), backend._stream())
            continue
        enc, data, high = tiles.encode(p, n, npdt)
        want = apply_tile_pass(want, n, p.low, high, enc, data, npdt)
        harr = (ctypes.c_int32 * max(1, len(high)))(*high)
        dbuf = ctypes.create_string_buffer(data, max(1, len(data)))
        _native.call("qj_apply_tile_pass", ctypes.c_void_p(state.data_ptr()), _code(dtype),
                     n, p.low, ctypes.cast(harr, ctypes.POINTER(ctypes.c_int32)), len(high),
                     enc.ctypes.data_as(ctypes.c_void_p), len(enc), ctypes.cast(dbuf, ctypes.c_void_p), len(data),
                     backend._stream())
    return backend.to_numpy(state), want


def _ops_with_swaps(n, rng, kmax):
    """Random circuit ops plus uncontrolled swaps (single and in runs) and a
    kmax-qubit dense block."""
    circuit = _random_circuit(n, 2, rng)
    out = []
    for i, g in enumerate(circuit):
        out.append(g)
        if i % 7 == 3:
            a, b = rng.choice(n, 2, replace=False)
            out.append(gates.SWAP(int(a), int(b)))
        if i % 11 == 5:
            q = rng.permutation(n)
            out += [gates.SWAP(int(q[0]), int(q[1])), gates.SWAP(int(q[2]), int(q[3]))]
    qs = [int(q) for q in rng.choice(n, kmax, replace=False)]
    out.insert(len(out) // 2, gates.Unitary(gm.haar_unitary(2**kmax, rng), *qs))
    return out


@pytest.mark.parametrize("kernel", ["register", "lds"])
@pytest.mark.parametrize("n,low,nhigh", [(10, 5, 5), (12, 6, 6), (11, 1, 7), (13, 4, 8), (14, 6, 7),
                                          (14, 3, 10), (13, 1, 11)])
def test_tile_kernel_matches_emulator(backend, monkeypatch, kernel, n, low, nhigh, dtype, rng):
    backend.set_dtype(dtype)
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    if (2 ** (low + nhigh)) * np.dtype(npdt).itemsize > 65536:
        pytest.skip("tile over 64 KiB")
    if kernel == "lds":
        monkeypatch.setenv("QJ_TILE_KERNEL", "lds")
    circuit = _ops_with_swaps(n, rng, 5 if dtype == "complex64" else 4)
    ops = [tiles.gate_to_op(port.custom_matrix, g, n) for g in circuit]
    ops = [op for op in ops if tiles.fits_alone(op, low, nhigh)]
    passes = tiles.plan(ops, n, low, nhigh)
    psi = gm.random_state(n, rng, npdt)
    got, want = _run_and_emulate(backend, passes, n, dtype, psi)
    np.testing.assert_allclose(got, want, atol=PARITY[dtype] * 10, rtol=0)


@pytest.mark.parametrize("nrb", ["3", "4", "5"])
def test_register_kernel_widths(backend, monkeypatch, nrb, dtype, rng):
    """QJ_TILE_NRB picks 2^nrb amplitudes per thread (4/5 complex64, 3/4
    complex128; other values fall back to the default)."""
    monkeypatch.setenv("QJ_TILE_NRB", nrb)
    backend.set_dtype(dtype)
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    n = 15
    low, nhigh = (6, 7) if dtype == "complex64" else (5, 7)
    circuit = _ops_with_swaps(n, rng, 3)
    ops = [tiles.gate_to_op(port.custom_matrix, g, n) for g in circuit]
    passes = tiles.plan(ops, n, low, nhigh)
    psi = gm.random_state(n, rng, npdt)
    got, want = _run_and_emulate(backend, passes, n, dtype, psi)
    np.testing.assert_allclose(got, want, atol=PARITY[dtype] * 10, rtol=0)


@pytest.mark.parametrize("n,pairs", [
    (12, [(0, 11), (1, 10), (2, 9), (3, 8), (4, 7), (5, 6)]),   # bit reversal
    (14, [(0, 3), (7, 12)]),                                      # low-low and high-high
    (9, [(2, 8)]),
    (16, [(6, 15), (7, 14), (0, 1)]),
    (5, [(0, 4), (1, 3)]),                                        # state smaller than a tile
])
def test_bit_swaps_match_emulator(backend, n, pairs, dtype, rng):
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    psi = gm.random_state(n, rng, npdt)
    got, want = _run_and_emulate(backend, [tiles.SwapPass(pairs)], n, dtype, psi)
    np.testing.assert_array_equal(got, want)


def test_bit_swaps_reject_bad_pairs(backend):
    state = backend.zero_state(8)
    for pairs in ([(0, 0)], [(0, 8)], [(1, 2), (2, 3)]):
        flat = (ctypes.c_int32 * (2 * len(pairs)))(*[b for pr in pairs for b in pr])
        with pytest.raises(ValueError):
            _native.call("qj_apply_bit_swaps", ctypes.c_void_p(state.data_ptr()), _native.QJ_COMPLEX128, 8,
                         ctypes.cast(flat, ctypes.POINTER(ctypes.c_int32)), len(pairs), backend._stream())


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


@pytest.mark.parametrize("kernel", ["register", "lds"])
def test_tile_pass_of_swaps_only(backend, monkeypatch, kernel, dtype, rng):
    """A pass whose only ops are in-tile swaps (pure relabelling: the store
    writes through the permuted bits)."""
    if kernel == "lds":
        monkeypatch.setenv("QJ_TILE_KERNEL", "lds")
    n = 14
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    low, nhigh = (6, 7) if dtype == "complex64" else (5, 7)
    p = tiles.Pass(low)
    p.high = [6, 8, 9, 10, 11, 12, 13]
    p.ops = [tiles.SwapOp(0, 13), tiles.SwapOp(2, 9), tiles.SwapOp(13, 6), tiles.SwapOp(1, 3)]
    psi = gm.random_state(n, rng, npdt)
    got, want = _run_and_emulate(backend, [p], n, dtype, psi)
    np.testing.assert_array_equal(got, want)


def test_tile_pass_with_plan_over_the_lds_budget(backend, rng):
    """Many phase ops (each carries per-segment factor tables) make the
    register kernel's plan too large for its LDS share: the pass runs on the
    LDS kernel and still matches the emulator."""
    n, dtype = 14, "complex128"
    p = tiles.Pass(5)
    p.high = list(range(5, 12))
    p.ops = []
    for i in range(30):
        c = int(rng.integers(0, 12))
        p.ops.append(tiles.PhaseOp({c}, 0.0, {b: 0.01 * (i + b) for b in range(n) if b != c}))
        p.ops.append(tiles.DenseOp([int(rng.integers(0, 12))], [], gm.haar_unitary(2, rng)))
    psi = gm.random_state(n, rng)
    got, want = _run_and_emulate(backend, [p], n, dtype, psi)
    np.testing.assert_allclose(got, want, atol=1e-11, rtol=0)
