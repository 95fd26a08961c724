"""Closed-form state-vector oracle (numpy) -- TEST INFRASTRUCTURE ONLY.

The state of n qubits is reshaped to (2,)*n with axis q = qubit q (qubit 0
is the most significant bit, src/qibojit/backends/cpu.py:565-569).  A gate
with matrix U on targets t_0..t_{k-1} (t_0 = matrix MSB) and controls c is
applied to the slice where every control axis is 1 by contracting U with the
target axes, which is how qibo's NumpyBackend applies gates and what the
reference's tests compare against (tests/test_gates.py).
"""

import numpy as np


def apply_gate(state, matrix, targets, controls=(), nqubits=None):
    """Return U applied to a copy of ``state`` (vector of 2^n amplitudes)."""
    state = np.asarray(state)
    if nqubits is None:
        nqubits = int(np.log2(state.size))
    k = len(targets)
    psi = state.reshape((2,) * nqubits).copy()
    u = np.asarray(matrix, dtype=np.complex128).reshape((2,) * (2 * k))
    sl = [slice(None)] * nqubits
    for c in controls:
        sl[c] = 1
    sl = tuple(sl)
    sub = psi[sl].astype(np.complex128)
    remaining = [a for a in range(nqubits) if a not in controls]
    t_axes = [remaining.index(t) for t in targets]
    res = np.tensordot(u, sub, axes=(list(range(k, 2 * k)), t_axes))
    res = np.moveaxis(res, list(range(k)), t_axes)
    psi[sl] = res
    return psi.reshape(-1).astype(state.dtype)


def apply_gate_density_matrix(rho, matrix, targets, controls=(), nqubits=None):
    """rho -> U rho U^dagger for a controlled gate."""
    rho = np.asarray(rho)
    d = rho.shape[0]
    if nqubits is None:
        nqubits = int(np.log2(d))
    full = full_matrix(matrix, targets, controls, nqubits)
    out = full @ rho.astype(np.complex128) @ full.conj().T
    return out.astype(rho.dtype)


def full_matrix(matrix, targets, controls, nqubits):
    """Dense 2^n x 2^n operator, column by column through ``apply_gate``."""
    d = 2**nqubits
    cols = [apply_gate(np.eye(d, dtype=np.complex128)[:, j], matrix, targets, controls, nqubits) for j in range(d)]
    return np.stack(cols, axis=1)


def kron_operator(matrix, targets, controls, nqubits):
    """Brute force by Kronecker products: sum over control patterns.

    Independent of ``apply_gate``: the gate is permuted to act on the
    leading qubits with np.kron, then the qubit axes are permuted back.
    """
    targets, controls = list(targets), list(controls)
    k, c = len(targets), len(controls)
    rest = [q for q in range(nqubits) if q not in targets and q not in controls]
    order = controls + targets + rest  # new axis i holds old qubit order[i]
    u = np.asarray(matrix, dtype=np.complex128)
    # controlled-U on (controls, targets): identity except all-ones block
    cu = np.eye(2 ** (c + k), dtype=np.complex128)
    cu[-(2**k):, -(2**k):] = u
    op = np.kron(cu, np.eye(2 ** len(rest)))
    # permute from `order` layout back to natural qubit order
    n = nqubits
    op = op.reshape((2,) * (2 * n))
    inv = [order.index(q) for q in range(n)]
    op = op.transpose(inv + [n + i for i in inv])
    return op.reshape(2**n, 2**n)


def collapse(state, qubits, results, normalize=True):
    """Zero every amplitude whose measured qubits differ from ``results``
    (tests/test_ops.py:65-83)."""
    state = np.asarray(state)
    n = int(np.log2(state.size))
    psi = state.reshape((2,) * n)
    sl = [slice(None)] * n
    for q, r in zip(qubits, results):
        sl[q] = r
    sl = tuple(sl)
    out = np.zeros_like(psi)
    out[sl] = psi[sl]
    out = out.reshape(-1)
    if normalize:
        out = out / np.sqrt(np.sum(np.abs(out.astype(np.complex128)) ** 2))
    return out.astype(state.dtype)


def probabilities(state, qubits):
    """Marginal |psi|^2 over ``qubits``, first listed qubit = MSB of the index."""
    state = np.asarray(state, dtype=np.complex128)
    n = int(np.log2(state.size))
    p = (np.abs(state) ** 2).reshape((2,) * n)
    others = tuple(q for q in range(n) if q not in qubits)
    p = p.sum(axis=others)
    # remaining axes are in increasing qubit order: permute to list order
    remaining = sorted(qubits)
    p = p.transpose([remaining.index(q) for q in qubits])
    return p.reshape(-1)


def projector(nqubits, qubits, results):
    """Diagonal 0/1 projector onto ``qubits`` = ``results`` (qubit q = bit n-1-q)."""
    idx = np.arange(2**nqubits)
    keep = np.ones(idx.size, dtype=bool)
    for q, r in zip(qubits, results):
        keep &= ((idx >> (nqubits - 1 - q)) & 1) == r
    return np.diag(keep.astype(np.complex128))


def collapse_density_matrix(rho, qubits, results, normalize=True):
    """P rho P, divided by its trace when ``normalize`` (qibo's density-matrix
    collapse: every entry whose row or column disagrees with the outcome is
    zeroed)."""
    rho = np.asarray(rho, dtype=np.complex128)
    n = int(np.log2(rho.shape[0]))
    p = projector(n, qubits, results)
    out = p @ rho @ p
    if normalize:
        out = out / np.trace(out).real
    return out


def probabilities_density_matrix(rho, qubits):
    """Marginal of the diagonal of ``rho`` over ``qubits`` (first = MSB)."""
    rho = np.asarray(rho, dtype=np.complex128)
    n = int(np.log2(rho.shape[0]))
    p = np.real(np.diag(rho)).reshape((2,) * n)
    others = tuple(q for q in range(n) if q not in qubits)
    p = p.sum(axis=others)
    remaining = sorted(qubits)
    p = p.transpose([remaining.index(q) for q in qubits])
    return p.reshape(-1)


def tfim_matrix(nqubits, g=1.0):
    """Dense H = -sum Z_i Z_{i+1} - g sum X_i (open chain)."""
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Z = np.array([[1, 0], [0, -1]], dtype=complex)

    def op_at(ops):
        m = np.array([[1.0 + 0j]])
        for q in range(nqubits):
            m = np.kron(m, ops.get(q, np.eye(2)))
        return m

    h = np.zeros((2**nqubits, 2**nqubits), dtype=complex)
    for i in range(nqubits - 1):
        h -= op_at({i: Z, i + 1: Z})
    for i in range(nqubits):
        h -= g * op_at({i: X})
    return h


def tfim_energy(state, g=1.0):
    """<psi|H|psi> by the closed form of the diagonal and X terms."""
    psi = np.asarray(state, dtype=np.complex128)
    n = int(np.log2(psi.size))
    idx = np.arange(psi.size)
    bits = (idx[:, None] >> (n - 1 - np.arange(n))[None, :]) & 1  # bit of qubit q
    z = 1 - 2 * bits
    diag = -np.sum(z[:, :-1] * z[:, 1:], axis=1)
    e = np.sum(diag * np.abs(psi) ** 2)
    for q in range(n):
        flip = idx ^ (1 << (n - 1 - q))
        e -= g * np.real(np.vdot(psi, psi[flip]))
    return float(np.real(e))
