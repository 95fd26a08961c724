"""Circuits with measurement gates through HipBackend.execute_circuit.

The reference builds a CircuitResult when the circuit measures
(src/qibojit/backends/gpu.py:722-729) and samples through sample_shots /
sample_frequencies (gpu.py:745-778, cpu.py:383-394); its tests run measured
circuits at tests/test_gates.py:369-411.  Expected outcomes here are drawn
with numpy.random.RandomState from the same seed (the generator behind the
reference's host sampling), and Metropolis frequencies come from the CPU
port of the reference's chains (ops.py:82-108).
"""

from collections import Counter

import numpy as np
import pytest

from oracle import einsum_ref as er
from oracle import gate_matrices as gm
from oracle import port

pytestmark = pytest.mark.gpu


def _circuit(n, rng, measured):
    from qibojit_amd._qibo import Circuit, gates

    c = Circuit(n)
    for q in range(n):
        c.add(gates.H(q))
        c.add(gates.U3(q, *rng.uniform(0, 2 * np.pi, 3)))
    for q in range(n - 1):
        c.add(gates.CNOT(q, q + 1))
    c.add(gates.fSim(0, n - 1, 0.4, 0.9))
    for group in measured:
        c.add(gates.M(*group))
    return c


def _oracle_state(port_backend, circuit):
    """Final state of the circuit's unitary part on the CPU port."""
    from qibojit_amd._qibo import Circuit

    unitary = Circuit(circuit.nqubits)
    for g in circuit.queue:
        if g.__class__.__name__ != "M":
            unitary.add(g)
    return port_backend.execute_circuit(unitary)


@pytest.mark.parametrize("fusion", [True, False])
def test_measured_circuit_1000_shots(backend, port_backend, fusion):
    """Frequencies and samples of a measured circuit below the Metropolis
    threshold: probabilities of the measured qubits (gate order, first qubit
    most significant) equal the closed form, and the shots equal
    RandomState(seed).choice(2^k, nshots, p)."""
    rng = np.random.default_rng(5)
    c = _circuit(5, rng, [(0, 2), (3,)])
    backend.fusion = fusion
    try:
        backend.set_seed(123)
        res = backend.execute_circuit(c, nshots=1000)
    finally:
        backend.fusion = True
    probs = backend.to_numpy(res.probabilities())
    want_probs = er.probabilities(_oracle_state(port_backend, c), [0, 2, 3])
    np.testing.assert_allclose(probs, want_probs, atol=1e-12)
    freqs = res.frequencies(binary=False)
    shots = np.random.RandomState(123).choice(8, size=1000, p=probs)
    assert dict(freqs) == {int(k): int(v) for k, v in zip(*np.unique(shots, return_counts=True))}
    binary = res.frequencies(binary=True)
    assert dict(binary) == {format(int(k), "03b"): int(v) for k, v in freqs.items()}
    assert sum(binary.values()) == 1000

    # samples of a fresh result are the same draws, as 0/1 rows
    backend.set_seed(123)
    res2 = backend.execute_circuit(c, nshots=1000)
    samples = res2.samples()
    assert samples.shape == (1000, 3)
    np.testing.assert_array_equal(res2.samples(binary=False), shots)
    np.testing.assert_array_equal(samples, (shots[:, None] >> np.array([2, 1, 0])) & 1)
    # frequencies after samples count those samples
    assert dict(res2.frequencies(binary=False)) == dict(freqs)


def test_measured_circuit_200000_shots_metropolis(backend, port_backend):
    """From SHOT_METROPOLIS_THRESHOLD shots on, frequencies come from the
    Metropolis chains: the chain seed is randint(0, 1e8) from the backend's
    generator (cpu.py:387) and the counts equal the port's chains bit for
    bit on the same probabilities."""
    from qibojit_amd._qibo import SHOT_METROPOLIS_THRESHOLD

    nshots = 200000
    assert nshots >= SHOT_METROPOLIS_THRESHOLD
    rng = np.random.default_rng(6)
    c = _circuit(6, rng, [(1, 4, 5)])
    backend.set_seed(2024)
    res = backend.execute_circuit(c, nshots=nshots)
    probs = backend.to_numpy(res.probabilities())
    np.testing.assert_allclose(probs, er.probabilities(_oracle_state(port_backend, c), [1, 4, 5]), atol=1e-12)
    freqs = res.frequencies(binary=False)
    seed = np.random.RandomState(2024).randint(0, int(1e8), size=1, dtype=np.int64)[0]
    want = port.measure_frequencies(np.zeros(8, dtype=np.int64), probs, nshots, 3, seed, backend.nthreads)
    assert dict(freqs) == {i: int(f) for i, f in enumerate(want) if f > 0}
    assert sum(freqs.values()) == nshots


def test_collapsing_measurement_runs_per_shot(backend):
    """A mid-circuit M(collapse=True) makes qibo run the circuit once per
    shot: each run draws the collapsing outcome, projects, continues, and
    one final shot of every measured qubit (0 from the collapsing gate, then
    1 and 2) is sampled from the final state.  The expected shots
    replay that sequence of draws with RandomState on the closed-form
    states."""
    from qibojit_amd._qibo import Circuit, gates

    theta = 0.7
    c = Circuit(3)
    c.add(gates.H(0))
    c.add(gates.CNOT(0, 1))
    c.add(gates.RY(2, theta))
    c.add(gates.M(0, collapse=True))
    c.add(gates.H(0))
    c.add(gates.CU3(1, 2, 0.3, 0.2, 0.9))
    c.add(gates.M(1, 2))
    nshots = 200
    backend.set_seed(77)
    res = backend.execute_circuit(c, nshots=nshots)
    got = np.asarray(res.samples())
    assert got.shape == (nshots, 3)

    rs = np.random.RandomState(77)
    want, collapsed = [], []
    for _ in range(nshots):
        psi = np.zeros(8, dtype=np.complex128)
        psi[0] = 1
        psi = er.apply_gate(psi, gm.H, [0])
        psi = er.apply_gate(psi, gm.X, [1], [0])
        psi = er.apply_gate(psi, gm.RY(theta), [2])
        p0 = er.probabilities(psi, [0])
        outcome = int(rs.choice(2, size=1, p=p0)[0])
        collapsed.append(outcome)
        psi = er.collapse(psi, [0], [outcome])
        psi = er.apply_gate(psi, gm.H, [0])
        psi = er.apply_gate(psi, gm.U3(0.3, 0.2, 0.9), [2], [1])
        p = er.probabilities(psi, [0, 1, 2])
        want.append(int(rs.choice(8, size=1, p=p / p.sum())[0]))
    want = np.array(want)
    np.testing.assert_array_equal(np.asarray(res.samples(binary=False)), want)
    assert dict(res.frequencies(binary=False)) == dict(Counter(int(w) for w in want))
    # qubit 1 keeps the collapsed outcome of qubit 0 (CNOT before the
    # collapse); both outcomes occur
    np.testing.assert_array_equal(got[:, 1], collapsed)
    assert set(collapsed) == {0, 1}


def test_circuit_without_measurements_returns_state(backend):
    from qibojit_amd._qibo import Circuit, gates

    c = Circuit(3)
    c.add(gates.H(1))
    res = backend.execute_circuit(c, nshots=10)
    assert not hasattr(res, "frequencies")
    want = np.zeros(8)
    want[[0, 2]] = 2**-0.5
    np.testing.assert_allclose(backend.to_numpy(res.state()), want, atol=1e-15)


def test_state_too_large_for_the_device(backend):
    """The allocation is checked against the device's capacity first:
    zero_state raises MemoryError and execute_circuit RuntimeError, as
    qibo's execute_circuit maps the backend's out-of-memory error."""
    from qibojit_amd._qibo import Circuit, gates

    with pytest.raises(MemoryError):
        backend.zero_state(40)
    c = Circuit(40)
    c.add(gates.H(0))
    with pytest.raises(RuntimeError, match="does not fit"):
        backend.execute_circuit(c)
