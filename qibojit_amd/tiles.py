"""Planner of fused tile passes (the gate-fusion executor).

A circuit segment of unitary gates is cut into passes.  Each pass streams
the state once through LDS tiles (``qj_apply_tile_pass``) whose bits are
the ``low`` lowest bit positions plus up to ``max_high`` chosen high
positions.  Gates become ops of two kinds:

* dense: the target-block matrix (controls kept as masks); its targets must
  be tile bits, so a dense gate on new high bits may start a new pass;
* phase: a diagonal gate on one target whose entries have modulus 1 (Z, S,
  T, RZ, U1, CZ, CU1, CRZ, CCZ, ...): amplitude i is multiplied by
  exp(i*(c0 + sum_j theta_j bit_j(i))) under its controls.  Phase ops fit
  any pass wherever their bits are, and consecutive ones merge: same
  control set -> angles add; a controlled phase with one control and no
  constant is symmetric in its two bits, so a ladder CU1(j, i), j > i, of
  a QFT merges into one op controlled by i.

Passes keep the gates in circuit order, so the result equals gate-by-gate
execution up to rounding.
"""

import ctypes

import numpy as np

from qibojit_amd.fusion import target_block

TILE_DENSE, TILE_PHASE = 0, 1
MAX_K = 5
TILE_OP_DTYPE = np.dtype(
    [
        ("kind", "<i4"),
        ("k", "<i4"),
        ("target", "<i4", (MAX_K + 1,)),
        ("local_ctrl", "<u8"),
        ("global_ctrl", "<u8"),
        ("offset", "<u8"),
        ("reserved", "<u8"),
    ]
)
assert TILE_OP_DTYPE.itemsize == 64


class DenseOp:
    """Matrix on global target bits (first = matrix MSB) under control bits."""

    def __init__(self, targets, controls, matrix):
        self.targets = list(targets)
        self.controls = frozenset(controls)
        self.matrix = np.asarray(matrix, dtype=np.complex128)

    def bits(self):
        return set(self.targets) | set(self.controls)


class PhaseOp:
    """exp(i*(c0 + sum theta_b bit_b)) under control bits (global positions)."""

    def __init__(self, controls, c0, thetas):
        self.controls = frozenset(controls)
        self.c0 = float(c0)
        self.thetas = dict(thetas)

    def bits(self):
        return set(self.controls) | set(self.thetas)

    def symmetric(self):
        return len(self.controls) == 1 and self.c0 == 0.0 and len(self.thetas) == 1

    def swapped(self):
        (c,) = tuple(self.controls)
        ((t, th),) = tuple(self.thetas.items())
        return PhaseOp({t}, 0.0, {c: th})

    def merged(self, other):
        """Combined op if both act under the same controls, else None."""
        for a in (self, self.swapped()) if self.symmetric() else (self,):
            for b in (other, other.swapped()) if other.symmetric() else (other,):
                if a.controls == b.controls:
                    th = dict(a.thetas)
                    for bit, t in b.thetas.items():
                        th[bit] = th.get(bit, 0.0) + t
                    return PhaseOp(a.controls, a.c0 + b.c0, th)
        return None


def gate_to_op(custom_matrix_of, gate, nqubits):
    """Tile op of a unitary gate, or None when it cannot run in a tile."""
    if gate.name == "fanout":
        return None
    targets = list(gate.target_qubits)
    if len(targets) > MAX_K:
        return None
    u = target_block(custom_matrix_of, gate)
    tbits = [nqubits - 1 - q for q in targets]
    cbits = [nqubits - 1 - q for q in gate.control_qubits]
    if len(targets) == 1 and u[0, 1] == 0 and u[1, 0] == 0:
        d0, d1 = complex(u[0, 0]), complex(u[1, 1])
        if abs(abs(d0) - 1) < 1e-14 and abs(abs(d1) - 1) < 1e-14:
            c0 = float(np.angle(d0))
            return PhaseOp(cbits, c0, {tbits[0]: float(np.angle(d1)) - c0})
    return DenseOp(tbits, cbits, u)


def column_op(op, nqubits):
    """Column half of rho -> U rho U^dagger on the flattened density matrix.

    ``op`` acts on the row bits (qubit q -> bit 2n-1-q); the same gate acts
    on the column bits (bit n-1-q) with the complex-conjugate matrix
    (src/qibojit/backends/cpu.py:474-496).
    """
    if isinstance(op, DenseOp):
        return DenseOp([b - nqubits for b in op.targets], [c - nqubits for c in op.controls], np.conj(op.matrix))
    return PhaseOp([c - nqubits for c in op.controls], -op.c0, {b - nqubits: -t for b, t in op.thetas.items()})


def expand_fanout(gate):
    """FanOut as its CNOTs (src/qibojit/backends/cpu.py:417-431)."""
    from qibojit_amd._qibo import gates  # pylint: disable=import-outside-toplevel

    control = gate.control_qubits[0]
    return [gates.CNOT(control, t) for t in gate.target_qubits]


class Pass:
    def __init__(self, low):
        self.low = low
        self.high = []
        self.ops = []

    def fits(self, op, max_high):
        if isinstance(op, PhaseOp):
            return True
        need = {b for b in op.targets if b >= self.low} - set(self.high)
        return len(self.high) + len(need) <= max_high

    def add(self, op):
        if isinstance(op, DenseOp):
            for b in op.targets:
                if b >= self.low and b not in self.high:
                    self.high.append(b)
        elif self.ops and isinstance(self.ops[-1], PhaseOp):
            m = self.ops[-1].merged(op)
            if m is not None:
                self.ops[-1] = m
                return
        self.ops.append(op)


def fits_alone(op, low, max_high):
    """Whether ``op`` fits an empty pass of this geometry."""
    return Pass(low).fits(op, max_high)


def plan(ops, nqubits, low, max_high):
    """Greedy cut of an op list into passes (circuit order preserved); every
    op must satisfy ``fits_alone``."""
    passes = []
    cur = Pass(low)
    for op in ops:
        if not cur.fits(op, max_high):
            passes.append(cur)
            cur = Pass(low)
        cur.add(op)
    if cur.ops:
        passes.append(cur)
    for p in passes:
        # widen the tile to the full budget with the lowest unused bits
        free = [b for b in range(low, nqubits) if b not in p.high]
        p.high = sorted(p.high + free[: max(0, min(max_high, nqubits - low) - len(p.high))])
    return passes


PHASE_LA = 8  # tile-local bits indexing phase table A (kernel constant)


def encode(p, nqubits, npdtype):
    """(ops array, data bytes, high bits) of a pass for qj_apply_tile_pass."""
    low, high = p.low, p.high
    tb = low + len(high)
    la = min(PHASE_LA, tb)
    local = {b: b for b in range(low)}
    local.update({b: low + j for j, b in enumerate(high)})
    # out-of-tile bits, ascending: bit r of the tile index t is outside[r]
    outside = [b for b in range(nqubits) if b not in local]
    nout = len(outside)
    ga = (nout + 1) // 2
    rank = {b: r for r, b in enumerate(outside)}
    ops = np.zeros(len(p.ops), dtype=TILE_OP_DTYPE)
    chunks, off = [], 0

    def put(buf):
        nonlocal off
        pad = (-len(buf)) % 16
        chunks.append(buf + b"\0" * pad)
        start = off
        off += len(buf) + pad
        return start

    for i, op in enumerate(p.ops):
        lmask = sum(1 << local[c] for c in op.controls if c in local)
        gmask = sum(1 << c for c in op.controls if c not in local)
        ops[i]["local_ctrl"] = lmask
        ops[i]["global_ctrl"] = gmask
        if isinstance(op, DenseOp):
            ops[i]["kind"] = TILE_DENSE
            ops[i]["k"] = len(op.targets)
            ops[i]["target"][: len(op.targets)] = [local[b] for b in op.targets]
            ops[i]["offset"] = put(np.ascontiguousarray(op.matrix, dtype=npdtype).tobytes())
        else:
            ops[i]["kind"] = TILE_PHASE
            head = np.array([ga, nout - ga, la, 0], dtype="<i4")
            # angle tables: OA/OB over the tile index t (c0 folded into OA),
            # A/B over the tile-local bits below / from la
            tabs = [np.zeros(2**ga), np.zeros(2 ** (nout - ga)), np.zeros(2**la), np.zeros(2 ** (tb - la))]
            tabs[0] += op.c0
            for b, th in op.thetas.items():
                if b in local:
                    lb = local[b]
                    k, sh = (2, lb) if lb < la else (3, lb - la)
                else:
                    r = rank[b]
                    k, sh = (0, r) if r < ga else (1, r - ga)
                idx = np.arange(tabs[k].size)
                tabs[k] += th * ((idx >> sh) & 1)
            blob = head.tobytes() + b"".join(np.exp(1j * a).astype(npdtype).tobytes() for a in tabs)
            ops[i]["offset"] = put(blob)
    data = b"".join(chunks)
    return ops, data, high


def run_passes(lib_call, state_ptr, dtype_code, nqubits, passes, npdtype, stream):
    """Issue one qj_apply_tile_pass per pass (stream-ordered, no host sync)."""
    for p in passes:
        ops, data, high = encode(p, nqubits, npdtype)
        harr = (ctypes.c_int32 * max(1, len(high)))(*high)
        dbuf = ctypes.create_string_buffer(data, max(1, len(data)))
        lib_call(
            "qj_apply_tile_pass",
            state_ptr,
            dtype_code,
            nqubits,
            p.low,
            ctypes.cast(harr, ctypes.POINTER(ctypes.c_int32)),
            len(high),
            ops.ctypes.data_as(ctypes.c_void_p),
            len(ops),
            ctypes.cast(dbuf, ctypes.c_void_p),
            len(data),
            stream,
        )
