"""Host-side matrices of fused gate blocks.

A FusedGate (qibo ``Circuit.fuse``) is applied as one dense gate over its
sorted qubits; its matrix is the ordered product of its gates, each embedded
in the block's qubits with its controls.  This is what qibo's
``Backend.matrix_fused`` provides to ``_as_custom_matrix``
(src/qibojit/backends/cpu.py:535-537, gpu.py:929-930).  Matrices are small
(at most 2^k x 2^k with k the fusion width), so they are built in complex128
with numpy on the host.
"""

import numpy as np


# Kernel family of a gate class (src/qibojit/backends/cpu.py:23-36).
GATE_OPS = {
    "X": "apply_x",
    "CNOT": "apply_x",
    "TOFFOLI": "apply_x",
    "Y": "apply_y",
    "Z": "apply_z",
    "CY": "apply_y",
    "CZ": "apply_z",
    "U1": "apply_z_pow",
    "CU1": "apply_z_pow",
    "SWAP": "apply_swap",
    "fSim": "apply_fsim",
    "GeneralizedfSim": "apply_fsim",
}


def target_block(custom_matrix_of, gate):
    """Dense target-block matrix of ``gate`` (controls excluded), from the
    kernel-facing encoding ``custom_matrix_of(gate)``."""
    name = gate.__class__.__name__
    custom = custom_matrix_of(gate)
    op = GATE_OPS.get(name)
    if op == "apply_x":
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if op == "apply_y":
        return np.array([[0, -1j], [1j, 0]], dtype=complex)
    if op == "apply_z":
        return np.diag([1, -1]).astype(complex)
    if op == "apply_z_pow":
        return np.diag([1, complex(np.asarray(custom))])
    if op == "apply_swap":
        return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)
    if op == "apply_fsim":
        c = np.asarray(custom, dtype=complex).reshape(-1)
        m = np.eye(4, dtype=complex)
        m[1, 1], m[1, 2], m[2, 1], m[2, 2] = c[0], c[1], c[2], c[3]
        m[3, 3] = c[4]
        return m
    return np.asarray(custom, dtype=complex)


def embed(block_qubits, targets, controls, u):
    """2^k x 2^k matrix of ``u`` on ``targets`` controlled by ``controls``,
    over ``block_qubits`` (first = most significant bit)."""
    k = len(block_qubits)
    pos = {q: i for i, q in enumerate(block_qubits)}
    t_axes = [pos[q] for q in targets]
    c_axes = [pos[q] for q in controls]
    nt = len(targets)
    full = np.eye(2**k, dtype=complex).reshape((2,) * k + (2**k,))
    u = np.asarray(u, dtype=complex).reshape((2,) * (2 * nt))
    # slice where all controls are 1
    sl = [slice(None)] * (k + 1)
    for a in c_axes:
        sl[a] = 1
    sl = tuple(sl)
    sub = full[sl]
    # remaining axes of `sub` after removing control axes
    remaining = [a for a in range(k) if a not in c_axes]
    sub_t_axes = [remaining.index(a) for a in t_axes]
    res = np.tensordot(u, sub, axes=(list(range(nt, 2 * nt)), sub_t_axes))
    # tensordot puts u's output axes first: move them back to their positions
    res = np.moveaxis(res, list(range(nt)), sub_t_axes)
    full[sl] = res
    return full.reshape(2**k, 2**k)


def fused_matrix(custom_matrix_of, fgate):
    """Ordered product of the block's gates over ``fgate.target_qubits``."""
    block = list(fgate.target_qubits)
    k = len(block)
    m = np.eye(2**k, dtype=complex)
    for gate in fgate.gates:
        u = target_block(custom_matrix_of, gate)
        g = embed(block, gate.target_qubits, gate.control_qubits, u)
        m = g @ m
    return m
