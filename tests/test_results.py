"""Measurement results and repeated execution of the qibo stand-in, on the
CPU with a NumPy backend built on the einsum oracle (host logic only; the
device path is tests/test_measurements_gpu.py)."""

from collections import Counter

import numpy as np
import pytest

from oracle import einsum_ref as er
from oracle import gate_matrices as gm
from qibojit_amd.qibo_compat import gates
from qibojit_amd.qibo_compat.backend import Backend
from qibojit_amd.qibo_compat.models import (
    Circuit,
    CircuitResult,
    MeasurementOutcomes,
    frequencies_to_binary,
)
from tests import product_states as ps


class EinsumBackend(Backend):
    """State vectors as numpy arrays, gates by the einsum closed form."""

    def zero_state(self, nqubits, density_matrix=False):
        state = np.zeros(2**nqubits, dtype=complex)
        state[0] = 1
        return state

    def cast(self, array, dtype=None, copy=False):
        return np.array(array, dtype=complex, copy=True)

    def nqubits_of(self, state):
        return int(np.log2(len(state)))

    def apply_gate(self, gate, state, nqubits):
        return er.apply_gate(state, ps.math_matrix(gate), gate.target_qubits, gate.control_qubits, nqubits)

    def calculate_probabilities(self, state, qubits, nqubits):
        return er.probabilities(state, list(qubits))

    def collapse_state(self, state, qubits, shot, nqubits, normalize=True):
        k = len(qubits)
        results = [(int(shot) >> (k - 1 - j)) & 1 for j in range(k)]
        return er.collapse(state, list(qubits), results, normalize)


def test_sample_conversions_round_trip():
    b = EinsumBackend()
    values = np.array([0, 5, 7, 2, 4])
    binary = b.samples_to_binary(values, 3)
    np.testing.assert_array_equal(binary, [[0, 0, 0], [1, 0, 1], [1, 1, 1], [0, 1, 0], [1, 0, 0]])
    np.testing.assert_array_equal(b.samples_to_decimal(binary, 3), values)
    assert frequencies_to_binary(Counter({5: 3, 0: 1}), 3) == Counter({"101": 3, "000": 1})


def test_outcomes_need_measurements_and_data():
    b = EinsumBackend()
    with pytest.raises(ValueError):
        MeasurementOutcomes([], b, probabilities=np.ones(1))
    with pytest.raises(ValueError):
        MeasurementOutcomes([gates.M(0)], b)


def test_circuit_result_draws_like_numpy_choice():
    """Frequencies come from sample_frequencies on the marginal of the
    measured qubits in gate order; samples from sample_shots; once samples
    exist, frequencies count them."""
    b = EinsumBackend()
    c = Circuit(4)
    for q, t in enumerate([0.3, 1.1, 2.0, 0.7]):
        c.add(gates.RY(q, t))
    c.add(gates.CNOT(0, 3))
    c.add(gates.M(3, 1))
    c.add(gates.M(0))
    state = b.zero_state(4)
    for g in c.queue:
        if not isinstance(g, gates.M):
            state = b.apply_gate(g, state, 4)
    b.set_seed(11)
    res = CircuitResult(state, c.measurements, b, nshots=500)
    probs = er.probabilities(state, [3, 1, 0])
    np.testing.assert_allclose(res.probabilities(), probs, atol=1e-15)
    shots = np.random.RandomState(11).choice(8, size=500, p=probs / probs.sum())
    assert dict(res.frequencies(binary=False)) == dict(Counter(int(s) for s in shots))
    b.set_seed(11)
    res2 = CircuitResult(state, c.measurements, b, nshots=500)
    np.testing.assert_array_equal(res2.samples(binary=False), shots)
    assert res2.samples().shape == (500, 3)
    assert dict(res2.frequencies(binary=False)) == dict(res.frequencies(binary=False))


def test_samples_after_frequencies_realise_them():
    b = EinsumBackend()
    state = np.array([0.6, 0.0, 0.0, 0.8], dtype=complex)
    b.set_seed(3)
    res = CircuitResult(state, [gates.M(0, 1)], b, nshots=100)
    freqs = res.frequencies(binary=False)
    values = res.samples(binary=False)
    assert dict(Counter(int(v) for v in values)) == dict(freqs)
    assert set(freqs) <= {0, 3}


def test_repeated_execution_replays_the_draws():
    """A collapsing measurement: one run per shot, the collapsing outcome
    drawn first, then one shot of every measured qubit of the final state."""
    b = EinsumBackend()
    c = Circuit(2)
    c.add(gates.H(0))
    c.add(gates.M(0, collapse=True))
    c.add(gates.CNOT(0, 1))
    c.add(gates.RY(0, 0.9))
    c.add(gates.M(1))
    assert c.repeated_execution
    b.set_seed(5)
    res = b.execute_circuit_repeated(c, 60)
    rs = np.random.RandomState(5)
    want = []
    for _ in range(60):
        psi = er.apply_gate(np.array([1, 0, 0, 0], dtype=complex), gm.H, [0])
        o = int(rs.choice(2, size=1, p=er.probabilities(psi, [0]))[0])
        psi = er.collapse(psi, [0], [o])
        psi = er.apply_gate(psi, gm.X, [1], [0])
        psi = er.apply_gate(psi, gm.RY(0.9), [0])
        p = er.probabilities(psi, [0, 1])
        want.append(int(rs.choice(4, size=1, p=p / p.sum())[0]))
    np.testing.assert_array_equal(res.samples(binary=False), want)
    assert sum(res.frequencies().values()) == 60
    # qubit 1 copies the collapsed outcome: both values appear
    assert set(res.samples()[:, 1]) == {0, 1}
