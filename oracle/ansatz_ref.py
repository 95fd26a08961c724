"""Hardware-efficient-ansatz oracle (numpy, complex128) -- TEST INFRASTRUCTURE ONLY.

The ansatz has no counterpart in the reference; it is built from qibo's
RY/RZ/CNOT gates applied the way the reference's tests apply gates
(einsum_ref.apply_gate, tests/test_gates.py) and scored with the TFIM energy
closed form (einsum_ref.tfim_energy).  Parameter (layer l, kind j, qubit q)
is at l*2n + j*n + q, j = 0 for RY and 1 for RZ.

Gradients come from two methods that share nothing with the adjoint method
under test:
  * parameter shift, dE/dt = (E(t + pi/2) - E(t - pi/2)) / 2, exact for
    gates exp(-i t P / 2) with a Pauli P;
  * central finite differences with step h (error O(h^2)).
Tests pin them against each other and against closed forms (E(0) = -(n-1)).
"""

import numpy as np

from oracle import einsum_ref
from oracle.gate_matrices import RY, RZ, X


def hea_state(nqubits, nlayers, params):
    """State after the ansatz, from |0...0>, as a complex128 vector."""
    p = np.asarray(params, dtype=np.float64).reshape(nlayers, 2, nqubits)
    psi = np.zeros(2**nqubits, dtype=np.complex128)
    psi[0] = 1.0
    for layer in range(nlayers):
        for q in range(nqubits):
            psi = einsum_ref.apply_gate(psi, RY(p[layer, 0, q]), [q], nqubits=nqubits)
        for q in range(nqubits):
            psi = einsum_ref.apply_gate(psi, RZ(p[layer, 1, q]), [q], nqubits=nqubits)
        for q in range(nqubits - 1):
            psi = einsum_ref.apply_gate(psi, X, [q + 1], controls=[q], nqubits=nqubits)
    return psi


def hea_energy(nqubits, nlayers, params, g=1.0):
    return einsum_ref.tfim_energy(hea_state(nqubits, nlayers, params), g)


def parameter_shift(nqubits, nlayers, params, g=1.0, indices=None):
    params = np.asarray(params, dtype=np.float64)
    indices = range(params.size) if indices is None else indices
    out = []
    for i in indices:
        plus, minus = params.copy(), params.copy()
        plus[i] += np.pi / 2
        minus[i] -= np.pi / 2
        out.append(0.5 * (hea_energy(nqubits, nlayers, plus, g) - hea_energy(nqubits, nlayers, minus, g)))
    return np.array(out)


def central_difference(nqubits, nlayers, params, g=1.0, h=1e-4, indices=None):
    params = np.asarray(params, dtype=np.float64)
    indices = range(params.size) if indices is None else indices
    out = []
    for i in indices:
        plus, minus = params.copy(), params.copy()
        plus[i] += h
        minus[i] -= h
        out.append((hea_energy(nqubits, nlayers, plus, g) - hea_energy(nqubits, nlayers, minus, g)) / (2 * h))
    return np.array(out)


def ladder_permutation(nqubits):
    """Index map of the CNOT ladder: psi_after[k] = psi_before[perm[k]]."""
    k = np.arange(2**nqubits)
    return k ^ (k >> 1)
