"""Gate-application parity on the GPU: HipBackend against the CPU port of
NumbaBackend (max |diff| <= 1e-12 complex128, 1e-5 complex64) and against
the einsum closed form (the reference's ATOL, tests/test_gates.py:15), on the
reference's case lists."""

import numpy as np
import pytest

from oracle import einsum_ref as er
from oracle import gate_matrices as gm
from oracle import port
from qibojit_amd._qibo import Circuit, gates
from tests import cases
from tests.conftest import ATOL, PARITY

pytestmark = pytest.mark.gpu


def _math_matrix(gate):
    name = gate.__class__.__name__
    p = gate.parameters
    table = {"X": gm.X, "Y": gm.Y, "Z": gm.Z, "H": gm.H, "CY": gm.Y, "CH": gm.H, "CSX": gm.SX,
             "CSXDG": gm.SXDG, "CZ": gm.Z, "CCZ": gm.Z, "CNOT": gm.X, "TOFFOLI": gm.X, "SWAP": gm.SWAP}
    if name in table:
        return table[name]
    if name == "Unitary":
        return np.asarray(p[0])
    if name in ("U1", "CU1"):
        return gm.U1(p[0])
    if name == "GeneralizedfSim":
        return gm.GeneralizedfSim(p[0], p[1])
    if name == "fSim":
        return gm.fSim(*p)
    return port.custom_matrix(gate)


def _check(backend, port_backend, gate, n, dtype, rng, density_matrix=False):
    backend.set_dtype(dtype)
    if density_matrix:
        state = gm.random_density_matrix(n, rng, dtype)
    else:
        state = gm.random_state(n, rng, dtype)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state, dtype=dtype, copy=True), n))
    want = port_backend.apply_gate(gate, state.copy(), n)
    assert got.dtype == np.dtype(dtype)
    np.testing.assert_allclose(got, want, atol=PARITY[dtype], rtol=0)
    u = _math_matrix(gate)
    if density_matrix:
        ref = er.apply_gate_density_matrix(state, u, gate.target_qubits, gate.control_qubits, n)
    else:
        ref = er.apply_gate(state, u, gate.target_qubits, gate.control_qubits, n)
    np.testing.assert_allclose(got, ref, atol=ATOL[dtype])


@pytest.mark.parametrize("n,target,controls", cases.APPLY_GATE)
def test_apply_gate(backend, port_backend, n, target, controls, dtype, rng):
    u = gm.haar_unitary(2, rng)
    _check(backend, port_backend, gates.Unitary(u, target).controlled_by(*controls), n, dtype, rng)


@pytest.mark.parametrize("n,target", cases.ONE_QUBIT_BASE)
@pytest.mark.parametrize("use_qubits", [False, True])
def test_one_qubit_base(backend, port_backend, n, target, use_qubits, dtype, rng):
    backend.set_dtype(dtype)
    state = gm.random_state(n, rng, dtype)
    matrix = (rng.random((2, 2)) + 1j * rng.random((2, 2))).astype(dtype)
    qubits = np.array([n - target - 1], dtype=np.int32) if use_qubits else None
    got = backend._one_qubit_base(backend.cast(state, dtype=dtype), n, target, "apply_gate",
                                  backend.cast(matrix, dtype=dtype), qubits)
    want = port_backend._one_qubit_base(state.copy(), n, target, "apply_gate", matrix, qubits)
    np.testing.assert_allclose(backend.to_numpy(got), want, atol=PARITY[dtype], rtol=0)


@pytest.mark.parametrize("n,target,controls", cases.PAULI)
@pytest.mark.parametrize("pauli", ["X", "Y", "Z"])
def test_apply_pauli_gate(backend, port_backend, n, target, pauli, controls, dtype, rng):
    _check(backend, port_backend, getattr(gates, pauli)(target).controlled_by(*controls), n, dtype, rng)


@pytest.mark.parametrize("n,target,controls", cases.ZPOW)
def test_apply_zpow_gate(backend, port_backend, n, target, controls, dtype, rng):
    _check(backend, port_backend, gates.U1(target, theta=0.1234).controlled_by(*controls), n, dtype, rng)


@pytest.mark.parametrize("n,targets,controls", cases.TWO_QUBIT)
@pytest.mark.parametrize("density_matrix", [False, True])
def test_apply_two_qubit_gate(backend, port_backend, n, targets, controls, density_matrix, dtype, rng):
    u = gm.haar_unitary(4, rng)
    _check(backend, port_backend, gates.Unitary(u, *targets).controlled_by(*controls), n, dtype, rng, density_matrix)


@pytest.mark.parametrize("n,targets", cases.TWO_QUBIT_BASE)
@pytest.mark.parametrize("use_qubits", [False, True])
def test_apply_two_qubit_base(backend, port_backend, n, targets, use_qubits, dtype, rng):
    backend.set_dtype(dtype)
    state = gm.random_state(n, rng, dtype)
    matrix = (rng.random((4, 4)) + 1j * rng.random((4, 4))).astype(dtype)
    qubits = np.array(sorted(n - q - 1 for q in targets), dtype=np.int32) if use_qubits else None
    got = backend._two_qubit_base(backend.cast(state, dtype=dtype), n, targets[0], targets[1],
                                  "apply_two_qubit_gate", matrix, qubits)
    want = port_backend._two_qubit_base(state.copy(), n, targets[0], targets[1], "apply_two_qubit_gate", matrix, qubits)
    np.testing.assert_allclose(backend.to_numpy(got), want, atol=PARITY[dtype], rtol=0)


@pytest.mark.parametrize("n,targets", cases.CONTROL_SINGLE)
@pytest.mark.parametrize("gatename", cases.CONTROL_SINGLE_GATES)
def test_apply_control_single_qubit_gate(backend, port_backend, gatename, n, targets, dtype, rng):
    _check(backend, port_backend, getattr(gates, gatename)(*targets), n, dtype, rng)


@pytest.mark.parametrize("n,target,controls", cases.CCZ)
def test_apply_ccz(backend, port_backend, n, target, controls, dtype, rng):
    _check(backend, port_backend, gates.CCZ(*controls, target), n, dtype, rng)


@pytest.mark.parametrize("n,targets", cases.DEUTSCH)
def test_apply_deutsch(backend, port_backend, n, targets, dtype, rng):
    _check(backend, port_backend, gates.DEUTSCH(*targets, float(rng.random())), n, dtype, rng)


@pytest.mark.parametrize("n,targets,controls", cases.SWAP)
def test_apply_swap(backend, port_backend, n, targets, controls, dtype, rng):
    _check(backend, port_backend, gates.SWAP(*targets).controlled_by(*controls), n, dtype, rng)


@pytest.mark.parametrize("n,targets,controls", cases.FSIM)
def test_apply_fsim(backend, port_backend, n, targets, controls, dtype, rng):
    u = rng.random((2, 2)) + 1j * rng.random((2, 2))
    _check(backend, port_backend, gates.GeneralizedfSim(*targets, u, cases.FSIM_PHI).controlled_by(*controls), n, dtype, rng)


@pytest.mark.parametrize("n", cases.FANOUT)
def test_apply_fanout(backend, port_backend, n, rng):
    gate = gates.FanOut(*range(n))
    state = gm.random_state(n, rng)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state, copy=True), n))
    want = port_backend.apply_gate(gate, state.copy(), n)
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)
    np.testing.assert_allclose(got, gate.matrix(backend) @ state, atol=1e-12)


@pytest.mark.parametrize("n,targets,controls", cases.MULTI_QUBIT)
@pytest.mark.parametrize("density_matrix", [False, True])
def test_apply_multiqubit_gate(backend, port_backend, n, targets, controls, density_matrix, dtype, rng):
    k = len(targets)
    u = rng.random((2**k, 2**k)) + 1j * rng.random((2**k, 2**k))
    if density_matrix and n > 6:
        # the dense closed form is O(4^n) memory: compare with the port only
        backend.set_dtype(dtype)
        rho = gm.random_density_matrix(n, rng, dtype)
        gate = gates.Unitary(u, *targets).controlled_by(*controls)
        got = backend.to_numpy(backend.apply_gate(gate, backend.cast(rho, dtype=dtype, copy=True), n))
        np.testing.assert_allclose(got, port_backend.apply_gate(gate, rho.copy(), n), atol=PARITY[dtype] * 4**k, rtol=0)
        return
    _check(backend, port_backend, gates.Unitary(u, *targets).controlled_by(*controls), n, dtype, rng, density_matrix)


@pytest.mark.parametrize("n,targets,controls", cases.MULTI_QUBIT_LARGE)
def test_apply_multiqubit_gate_large(backend, port_backend, n, targets, controls, dtype, rng):
    k = len(targets)
    u = gm.haar_unitary(2**k, rng)
    _check(backend, port_backend, gates.Unitary(u, *targets).controlled_by(*controls), n, dtype, rng)


@pytest.mark.parametrize("targets,controls", [([21, 20, 19, 18, 17], []), ([17, 21, 18, 20], [3]),
                                              ([19, 20, 21, 2, 9], [0, 16]), ([20, 5, 21, 8], [])])
def test_low_target_blocks(backend, port_backend, targets, controls, dtype, rng):
    """4/5 targets among the lowest bits (inside one cache line)."""
    k = len(targets)
    u = gm.haar_unitary(2**k, rng)
    _check(backend, port_backend, gates.Unitary(u, *targets).controlled_by(*controls), 22, dtype, rng)


@pytest.mark.parametrize("n,targets", cases.MULTI_QUBIT_BASE)
@pytest.mark.parametrize("use_qubits", [False, True])
def test_apply_multi_qubit_base(backend, port_backend, n, targets, use_qubits, dtype, rng):
    backend.set_dtype(dtype)
    state = gm.random_state(n, rng, dtype)
    matrix = (rng.random((8, 8)) + 1j * rng.random((8, 8))).astype(dtype)
    qubits = np.array(sorted(n - q - 1 for q in targets), dtype=np.int32) if use_qubits else None
    got = backend._multi_qubit_base(backend.cast(state, dtype=dtype), n, targets, matrix, qubits)
    want = port_backend._multi_qubit_base(state.copy(), n, targets, matrix, qubits)
    np.testing.assert_allclose(backend.to_numpy(got), want, atol=PARITY[dtype] * 8, rtol=0)


@pytest.mark.parametrize("k", [6, 7, 8, 10])
def test_wide_dense_gates(backend, port_backend, k, dtype, rng):
    """LDS-staged path for k > 5 targets (the reference's GPU limit is 7)."""
    n = k + 4
    targets = list(rng.permutation(n)[:k])
    u = gm.haar_unitary(2**k, rng)
    _check(backend, port_backend, gates.Unitary(u, *targets), n, dtype, rng)


def test_too_many_targets(backend):
    state = backend.zero_state(14)
    with pytest.raises(ValueError):
        backend._multi_qubit_base(state, 14, list(range(13)), np.eye(2**13), None)


@pytest.mark.parametrize("gatename", ["H", "X", "Y", "Z"])
@pytest.mark.parametrize("density_matrix", [False, True])
def test_gates_on_circuit(backend, port_backend, gatename, density_matrix, rng):
    state = gm.random_density_matrix(1, rng) if density_matrix else gm.random_state(1, rng)
    c = Circuit(1, density_matrix=density_matrix)
    c.add(getattr(gates, gatename)(0))
    got = backend.to_numpy(backend.execute_circuit(c, np.copy(state)).state())
    want = port_backend.apply_gate(c.queue[0], state.copy(), 1)
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)


@pytest.mark.parametrize("gatename,params", cases.PARAMETRIZED_ON_CIRCUIT)
@pytest.mark.parametrize("density_matrix", [False, True])
def test_parametrized_gates_on_circuit(backend, port_backend, gatename, params, density_matrix, rng):
    state = gm.random_density_matrix(2, rng) if density_matrix else gm.random_state(2, rng)
    c = Circuit(2, density_matrix=density_matrix)
    c.add(getattr(gates, gatename)(0, 1, **params))
    got = backend.to_numpy(backend.execute_circuit(c, np.copy(state)).state())
    want = port_backend.apply_gate(c.queue[0], state.copy(), 2)
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)


@pytest.mark.parametrize("gatename", ["H", "X", "Z"])
def test_density_matrix_half_calls(backend, gatename, rng):
    rho = gm.random_density_matrix(3, rng)
    gate = getattr(gates, gatename)(1)
    got = backend.to_numpy(backend.apply_gate_half_density_matrix(gate, backend.cast(rho, copy=True), 3))
    full = er.full_matrix(_math_matrix(gate), [1], [], 3)
    np.testing.assert_allclose(got, full @ rho, atol=1e-12)


def test_unitary_channel(backend, dtype, rng):
    backend.set_dtype(dtype)
    a1 = gm.X
    a2 = gm.controlled(gm.X)
    channel = gates.UnitaryChannel([(0,), (2, 3)], [(0.4, a1), (0.3, a2)])
    rho = gm.random_density_matrix(4, rng, dtype)
    got = backend.to_numpy(backend.apply_channel(channel, backend.cast(rho, dtype=dtype, copy=True), 4))
    u1 = er.full_matrix(a1, [0], [], 4)
    u2 = er.full_matrix(a2, [2, 3], [], 4)
    want = 0.3 * rho + 0.4 * u1 @ rho @ u1.conj().T + 0.3 * u2 @ rho @ u2.conj().T
    np.testing.assert_allclose(got, want, atol=ATOL[dtype])


def test_random_circuit_20q_complex128(backend, port_backend):
    """20-qubit depth-20 U3 + CZ-pairing circuit from |0...0>: <= 1e-12 vs the port."""
    n, depth = 20, 20
    rng = np.random.default_rng(1234)
    c = Circuit(n)
    for _ in range(depth):
        for q in range(n):
            c.add(gates.U3(q, *rng.uniform(0, 2 * np.pi, 3)))
        perm = rng.permutation(n)
        for j in range(n // 2):
            c.add(gates.CZ(int(perm[2 * j]), int(perm[2 * j + 1])))
    got = backend.to_numpy(backend.execute_circuit(c).state())
    want = port_backend.execute_circuit(c)
    assert np.abs(got - want).max() <= 1e-12
    assert abs(np.vdot(got, got).real - 1) < 1e-12


@pytest.mark.parametrize("width", [2, 3, 5])
def test_fused_circuit_matches_unfused(backend, port_backend, width, rng):
    n = 10
    c = Circuit(n)
    for layer in range(6):
        for q in range(n):
            c.add(gates.U3(q, *rng.uniform(0, 2 * np.pi, 3)))
        for q in range(layer % 2, n - 1, 2):
            c.add(gates.CNOT(q, q + 1))
        c.add(gates.fSim(1, 7, 0.3, 0.2))
        c.add(gates.CU1(8, 2, 0.4))
    psi0 = gm.random_state(n, rng)
    want = port_backend.execute_circuit(c, psi0)
    got = backend.to_numpy(backend.execute_circuit(c.fuse(max_qubits=width), psi0.copy()).state())
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)


def test_large_state_round_trip(backend, dtype):
    """28 qubits: U then U^dagger restores the state (size-independent check)."""
    backend.set_dtype(dtype)
    n = 28
    rng = np.random.default_rng(5)
    psi0 = gm.random_state(n, rng, dtype)
    state = backend.cast(psi0, dtype=dtype, copy=True)
    seq = []
    for t in (0, 13, 27):
        u = gm.haar_unitary(2, rng)
        seq.append(gates.Unitary(u, t))
        seq.append(gates.Unitary(gm.haar_unitary(4, rng), t, (t + 5) % n).controlled_by((t + 9) % n))
        seq.append(gates.Unitary(gm.haar_unitary(8, rng), t, (t + 1) % n, (t + 20) % n))
    for g in seq:
        state = backend.apply_gate(g, state, n)
    norm = backend.norm_squared(state)
    assert abs(norm - 1) < (1e-10 if dtype == "complex128" else 1e-4)
    for g in reversed(seq):
        m = np.asarray(g.parameters[0])
        inv = gates.Unitary(m.conj().T, *g.target_qubits).controlled_by(*g.control_qubits)
        state = backend.apply_gate(inv, state, n)
    overlap = backend.vdot(backend.cast(psi0, dtype=dtype), state)
    assert abs(overlap - 1) < (1e-10 if dtype == "complex128" else 1e-4)


def test_readout_error_channel(backend, rng):
    """tests/test_gates.py:443-455: a readout-error channel keeps the trace."""
    from qibojit_amd.qibo_compat.config import PRECISION_TOL
    from qibojit_amd.qibo_compat.gates import ReadoutErrorChannel

    rho = gm.random_density_matrix(1, rng)
    p = rng.random((2, 2))
    p /= p.sum(axis=1, keepdims=True)
    channel = ReadoutErrorChannel(0, p)
    out = backend.to_numpy(channel.apply(backend, backend.cast(rho.copy()), 1))
    assert abs(np.trace(out).real - 1) < PRECISION_TOL
    want = sum(k @ rho @ k.conj().T for k in (g.parameters[0] for g in channel.gates))
    np.testing.assert_allclose(out, want, atol=1e-12)


def test_kraus_channel(backend, dtype, rng):
    """rho -> sum K rho K^+ against the dense closed form."""
    from qibojit_amd.qibo_compat.gates import KrausChannel

    backend.set_dtype(dtype)
    n = 3
    rho = gm.random_density_matrix(n, rng)
    u1, u2 = gm.haar_unitary(4, rng), gm.haar_unitary(4, rng)
    ops = [np.sqrt(0.3) * u1, np.sqrt(0.7) * u2]
    channel = KrausChannel((0, 2), ops)
    out = backend.to_numpy(backend.apply_channel(channel, backend.cast(rho.copy()), n))
    full = [er.full_matrix(k, [0, 2], [], n) for k in ops]
    want = sum(f @ rho @ f.conj().T for f in full)
    np.testing.assert_allclose(out, want, atol=ATOL[dtype])
    assert abs(np.trace(out).real - 1) < (1e-5 if dtype == "complex64" else 1e-12)
