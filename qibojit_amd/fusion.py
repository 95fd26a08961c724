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


def plan_blocks(gate_qubits, max_qubits=5):
    """Group a gate sequence into blocks of at most ``max_qubits`` qubits.

    ``gate_qubits[i]`` is the qubit tuple (targets and controls) of gate i.
    A gate may join the open block only when every earlier gate on each of
    its qubits is already placed (in an earlier block or this one), so the
    blocks, applied in order, equal the sequence.  The block grows first by
    gates inside its qubits, then by the ready gate adding the fewest new
    qubits (earliest first); it closes when no ready gate fits.  Unlike
    qibo's ``Circuit.fuse`` (one owner block per qubit), blocks may take
    gates from several earlier chains, which is what lets a brickwork of
    2-qubit gates fuse into 5-qubit blocks.  Returns a list of
    (sorted block qubits, [gate indices in order]).
    """
    per_q = {}
    for i, qs in enumerate(gate_qubits):
        if len(set(qs)) > max_qubits:
            raise ValueError(f"Gate {i} acts on more than {max_qubits} qubits.")
        for q in qs:
            per_q.setdefault(q, []).append(i)
    head = {q: 0 for q in per_q}
    placed = [False] * len(gate_qubits)

    def ready(i):
        return all(per_q[q][head[q]] == i for q in gate_qubits[i])

    def place(i):
        placed[i] = True
        for q in gate_qubits[i]:
            head[q] += 1

    blocks = []
    nxt = 0
    while nxt < len(gate_qubits):
        while nxt < len(gate_qubits) and placed[nxt]:
            nxt += 1
        if nxt == len(gate_qubits):
            break
        block, bq = [nxt], set(gate_qubits[nxt])
        place(nxt)
        while True:
            cands = sorted({per_q[q][head[q]] for q in per_q if head[q] < len(per_q[q])})
            inside = [i for i in cands if set(gate_qubits[i]) <= bq and ready(i)]
            if inside:
                for i in inside:
                    if ready(i):
                        place(i)
                        block.append(i)
                continue
            best = None
            for i in cands:
                new = set(gate_qubits[i]) - bq
                if len(bq) + len(new) <= max_qubits and ready(i):
                    key = (len(new), i)
                    if best is None or key < best[0]:
                        best = (key, i)
            if best is None:
                break
            i = best[1]
            place(i)
            block.append(i)
            bq |= set(gate_qubits[i])
        blocks.append((sorted(bq), block))
    return blocks


def fuse_blocks(circuit, max_qubits=5):
    """A circuit whose unitary gates are grouped by :func:`plan_blocks` into
    FusedGates of at most ``max_qubits`` qubits (one dense pass each);
    single-gate blocks keep the gate itself.  Every gate must be unitary."""
    from qibojit_amd._qibo import Circuit, FusedGate  # pylint: disable=import-outside-toplevel

    queue = list(circuit.queue)
    blocks = plan_blocks([tuple(g.qubits) for g in queue], max_qubits)
    out = Circuit(circuit.nqubits)
    for qubits, idx in blocks:
        if len(idx) == 1:
            out.add(queue[idx[0]])
            continue
        fgate = FusedGate(*qubits)
        for i in idx:
            fgate.append(queue[i])
        out.add(fgate)
    return out
