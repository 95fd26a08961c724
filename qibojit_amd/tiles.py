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
  a QFT merges into one op controlled by i;
* swap: an uncontrolled SWAP, a relabelling of two tile bits inside a pass.

A run of consecutive SWAPs on distinct qubits that does not fit the current
pass becomes a bit-swap pass of its own (``qj_apply_bit_swaps``: one read
and one write for the whole run, e.g. the final swaps of a QFT).

Passes keep the gates in circuit order, so the result equals gate-by-gate
execution up to rounding.
"""

import cmath
import ctypes
import functools
import math
import struct

import numpy as np

from qibojit_amd.fusion import GATE_OPS, target_block

TILE_DENSE, TILE_PHASE, TILE_SWAP = 0, 1, 2
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
# one TILE_OP_DTYPE record packed in a single call (kind, k, target[6],
# local_ctrl, global_ctrl, offset, reserved); the phase-op data header
_OP_RECORD = struct.Struct("<ii6iQQQQ")
assert _OP_RECORD.size == TILE_OP_DTYPE.itemsize
_PHASE_HEAD = struct.Struct("<4i")
_ZEROS = (0,) * (MAX_K + 1)
_PAULI_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_PAULI_X.flags.writeable = False


class DenseOp:
    """Matrix on global target bits (first = matrix MSB) under control bits."""

    def __init__(self, targets, controls, matrix):
        self.targets = list(targets)
        self.controls = frozenset(controls)
        self.matrix = np.asarray(matrix, dtype=np.complex128)

    def bits(self):
        return set(self.targets) | set(self.controls)


class SwapOp:
    """Exchange of two global bits (an uncontrolled SWAP)."""

    def __init__(self, a, b):
        self.targets = [a, b]
        self.controls = frozenset()

    def bits(self):
        return set(self.targets)


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
    tbits = [nqubits - 1 - q for q in targets]
    cbits = [nqubits - 1 - q for q in gate.control_qubits]
    kind = GATE_OPS.get(type(gate).__name__)
    # fixed-matrix families skip the matrix encoding (small circuits spend
    # most of their host time here)
    if kind == "apply_z" and len(targets) == 1:
        return PhaseOp(cbits, 0.0, {tbits[0]: math.pi})  # diag(1, -1)
    if kind == "apply_x" and len(targets) == 1:
        return DenseOp(tbits, cbits, _PAULI_X)
    if kind == "apply_swap" and not cbits:
        return SwapOp(*tbits)
    if kind == "apply_z_pow" and len(targets) == 1:
        # U1 / CU1: diag(1, z) without building the matrix (the QFT's ladders
        # are hundreds of these)
        z = complex(np.asarray(custom_matrix_of(gate)))
        if abs(abs(z) - 1) < 1e-14:
            return PhaseOp(cbits, 0.0, {tbits[0]: cmath.phase(z)})
    u = target_block(custom_matrix_of, gate)
    if len(targets) == 1 and u[0, 1] == 0 and u[1, 0] == 0:
        d0, d1 = complex(u[0, 0]), complex(u[1, 1])
        if abs(abs(d0) - 1) < 1e-14 and abs(abs(d1) - 1) < 1e-14:
            c0 = cmath.phase(d0)
            return PhaseOp(cbits, c0, {tbits[0]: cmath.phase(d1) - c0})
    return DenseOp(tbits, cbits, u)


def column_op(op, nqubits):
    """Column half of rho -> U rho U^dagger on the flattened density matrix.

    ``op`` acts on the row bits (qubit q -> bit 2n-1-q); the same gate acts
    on the column bits (bit n-1-q) with the complex-conjugate matrix
    (src/qibojit/backends/cpu.py:474-496).
    """
    if isinstance(op, DenseOp):
        return DenseOp([b - nqubits for b in op.targets], [c - nqubits for c in op.controls], np.conj(op.matrix))
    if isinstance(op, SwapOp):
        return SwapOp(*(b - nqubits for b in op.targets))
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
        if isinstance(op, (DenseOp, SwapOp)):
            for b in op.targets:
                if b >= self.low and b not in self.high:
                    self.high.append(b)
        elif self.ops and isinstance(self.ops[-1], PhaseOp):
            m = self.ops[-1].merged(op)
            if m is not None:
                self.ops[-1] = m
                return
        self.ops.append(op)


class SwapPass:
    """Disjoint bit swaps applied in one pass (qj_apply_bit_swaps)."""

    def __init__(self, pairs):
        self.pairs = [tuple(p) for p in pairs]


def fits_alone(op, low, max_high):
    """Whether ``op`` fits an empty pass of this geometry."""
    return Pass(low).fits(op, max_high)


class Planner:
    """Greedy cut of an op stream into passes (circuit order preserved);
    every op must satisfy ``fits_alone``.  ``push(op)`` returns the passes
    that the op closed, ready to launch, so a caller can run a pass while
    it is still converting the gates of the next one; ``finish()`` returns
    the rest.  A run of two or more swaps on distinct bits that does not fit
    the current pass becomes a SwapPass (swaps are held back until the run
    ends)."""

    def __init__(self, nqubits, low, max_high):
        self.nqubits, self.low, self.max_high = nqubits, low, max_high
        self.cur = Pass(low)
        self.run = []  # pending SwapOps on distinct bits

    def _close(self, p):
        if isinstance(p, Pass):
            # widen the tile to the full budget with the lowest unused bits
            free = [b for b in range(self.low, self.nqubits) if b not in p.high]
            p.high = sorted(p.high + free[: max(0, min(self.max_high, self.nqubits - self.low) - len(p.high))])
        return p

    def _add(self, op, out):
        if not self.cur.fits(op, self.max_high):
            out.append(self._close(self.cur))
            self.cur = Pass(self.low)
        self.cur.add(op)

    def _flush_run(self, out):
        run, self.run = self.run, []
        if len(run) == 1:
            self._add(run[0], out)
        elif run:
            trial = Pass(self.low)
            trial.high = list(self.cur.high)
            fits = True
            for o in run:
                fits = fits and trial.fits(o, self.max_high)
                trial.add(o)
            if fits:
                for o in run:
                    self.cur.add(o)
            else:
                if self.cur.ops:
                    out.append(self._close(self.cur))
                out.append(SwapPass([o.targets for o in run]))
                self.cur = Pass(self.low)

    def push(self, op):
        out = []
        if isinstance(op, SwapOp):
            used = {b for o in self.run for b in o.targets}
            if used & set(op.targets):
                self._flush_run(out)
            self.run.append(op)
            return out
        self._flush_run(out)
        self._add(op, out)
        return out

    def finish(self):
        out = []
        self._flush_run(out)
        if self.cur.ops:
            out.append(self._close(self.cur))
        self.cur = Pass(self.low)
        return out


def plan(ops, nqubits, low, max_high):
    """The passes of a whole op list (see Planner)."""
    pl = Planner(nqubits, low, max_high)
    passes = []
    for op in ops:
        passes += pl.push(op)
    return passes + pl.finish()


@functools.lru_cache(maxsize=256)
def _bit_column(size, sh):
    """Bit ``sh`` of 0 .. size-1 as float64 (a read-only table)."""
    col = ((np.arange(size) >> sh) & 1).astype(np.float64)
    col.flags.writeable = False
    return col


def encode(p, nqubits, npdtype):
    """(ops array, data bytes, high bits) of a pass for qj_apply_tile_pass."""
    low, high = p.low, p.high
    tb = low + len(high)
    local = {b: b for b in range(low)}
    local.update({b: low + j for j, b in enumerate(high)})
    # out-of-tile bits, ascending: bit r of the tile index t is outside[r]
    outside = [b for b in range(nqubits) if b not in local]
    nout = len(outside)
    ga = (nout + 1) // 2
    rank = {b: r for r, b in enumerate(outside)}
    recs, chunks = [], []
    off = 0
    for op in p.ops:
        if isinstance(op, SwapOp):
            a, b = op.targets
            recs.append(_OP_RECORD.pack(TILE_SWAP, 2, local[a], local[b], 0, 0, 0, 0, 0, 0, 0, 0))
            continue
        lmask = gmask = 0
        for c in op.controls:
            if c in local:
                lmask |= 1 << local[c]
            else:
                gmask |= 1 << c
        if isinstance(op, DenseOp):
            tg = [local[b] for b in op.targets]
            blob = np.ascontiguousarray(op.matrix, dtype=npdtype).tobytes()
            recs.append(_OP_RECORD.pack(TILE_DENSE, len(tg), *tg, *_ZEROS[len(tg):], lmask, gmask, off, 0))
        else:
            # OA/OB: angle tables over the tile index t (c0 folded into OA);
            # W: one angle per tile-local bit
            oa, ob, w = np.full(2**ga, op.c0), np.zeros(2 ** (nout - ga)), np.zeros(tb)
            for b, th in op.thetas.items():
                if b in local:
                    w[local[b]] += th
                else:
                    r = rank[b]
                    tab, sh = (oa, r) if r < ga else (ob, r - ga)
                    tab += th * _bit_column(tab.size, sh)
            ang = np.concatenate((oa, ob, w))
            blob = _PHASE_HEAD.pack(ga, nout - ga, tb, 0) + np.exp(1j * ang).astype(npdtype).tobytes()
            recs.append(_OP_RECORD.pack(TILE_PHASE, 0, *_ZEROS, lmask, gmask, off, 0))
        pad = (-len(blob)) % 16
        chunks.append(blob + b"\0" * pad if pad else blob)
        off += len(blob) + pad
    ops = np.frombuffer(b"".join(recs), dtype=TILE_OP_DTYPE)
    return ops, b"".join(chunks), high


def run_passes(lib_call, state_ptr, dtype_code, nqubits, passes, npdtype, stream):
    """Issue one qj_apply_tile_pass (or qj_apply_bit_swaps) per pass
    (stream-ordered, no host sync)."""
    for p in passes:
        if isinstance(p, SwapPass):
            flat = [b for pair in p.pairs for b in pair]
            arr = (ctypes.c_int32 * len(flat))(*flat)
            lib_call("qj_apply_bit_swaps", state_ptr, dtype_code, nqubits,
                     ctypes.cast(arr, ctypes.POINTER(ctypes.c_int32)), len(p.pairs), stream)
            continue
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
