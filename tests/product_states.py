"""Closed form of a gate applied to a product state (test helper).

For a product state a_0 (x) a_1 (x) ... (x) a_{n-1} (qubit 0 the most
significant bit) and a gate U on qubits Q (controls first, then targets),
the amplitude at index i is

    prod_{q not in Q} a_q[i_q] * (U (x)_{q in Q} a_q)[i_Q]

with i_Q the bits of i on Q in that order.  Checked against the einsum
closed form in tests/test_oracle.py.
"""

import numpy as np

from oracle import gate_matrices as gm


def math_matrix(gate):
    """Oracle matrix over the gate's targets (controls not included)."""
    name = gate.__class__.__name__
    p = gate.parameters
    fixed = {"H": gm.H, "X": gm.X, "Y": gm.Y, "Z": gm.Z, "CNOT": gm.X, "SWAP": gm.SWAP}
    if name in fixed:
        return fixed[name]
    if name == "Unitary":
        return np.asarray(p[0], dtype=complex)
    if name in ("RX", "RY", "RZ"):
        return getattr(gm, name)(p[0])
    if name == "U1":
        return gm.U1(p[0])
    if name in ("U3", "CU3"):
        return gm.U3(*p)
    if name == "fSim":
        return gm.fSim(*p)
    if name == "GeneralizedfSim":
        return gm.GeneralizedfSim(p[0], p[1])
    raise NotImplementedError(name)


def factors(n, rng):
    """Random normalised one-qubit states a_q (rows), qubit 0 first."""
    a = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    return a / np.linalg.norm(a, axis=1, keepdims=True)


def device_product_state(factors, dtype):
    """a_0 (x) a_1 (x) ... on the device, built by broadcasting products
    (no kernel of the backend under test)."""
    import torch

    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    state = torch.tensor(factors[0], dtype=tdt, device="cuda")
    for f in factors[1:]:
        state = (state[:, None] * torch.tensor(f, dtype=tdt, device="cuda")[None, :]).reshape(-1)
    return state


def expected(factors, gate, idx):
    """Closed-form amplitudes at indices ``idx`` after ``gate`` on the
    product state of ``factors`` (FP64, from the stored-precision factors)."""
    n = len(factors)
    qs = list(gate.control_qubits) + list(gate.target_qubits)
    u = gm.controlled(math_matrix(gate), len(gate.control_qubits))
    bits = (idx[:, None] >> (n - 1 - np.arange(n))) & 1
    rest = np.ones(len(idx), dtype=complex)
    for q in range(n):
        if q not in qs:
            rest *= factors[q][bits[:, q]]
    local = np.ones(1, dtype=complex)
    for q in qs:
        local = np.kron(local, factors[q])
    w = u @ local
    sub = np.zeros(len(idx), dtype=np.int64)
    for q in qs:
        sub = (sub << 1) | bits[:, q]
    return rest * w[sub]


def host_product_state(factors):
    """The full product state on the host (small n)."""
    state = np.ones(1, dtype=complex)
    for f in factors:
        state = np.kron(state, f)
    return state
