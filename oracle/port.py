"""ctypes wrapper of numba_port.c and a NumbaBackend-shaped dispatcher.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline; see oracle/__init__.py).

``PortBackend`` restates NumbaBackend's host dispatch
(src/qibojit/backends/cpu.py:369-635) on numpy arrays, calling the C
restatements of the numba kernels.  Gate encodings follow CustomMatrices
(src/qibojit/backends/matrices.py:7-70) but are built from this package's
own textbook matrices.
"""

import ctypes
import os
import subprocess

import numpy as np

from oracle import gate_matrices as gm

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "numba_port.c")
LIB = os.path.join(HERE, "liboracle_port.so")

CFLAGS = ["-O3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11"]


def build():
    """Compile numba_port.c with gcc into liboracle_port.so."""
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", *CFLAGS, "-o", tmp, SRC, "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


_LIB = None


def lib():
    global _LIB  # pylint: disable=global-statement
    if _LIB is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        _LIB = ctypes.CDLL(LIB)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _suffix(dtype):
    return "c64" if np.dtype(dtype) == np.complex64 else "c128"


def set_threads(n):
    lib().port_set_threads(ctypes.c_int(int(n)))


def max_threads():
    return int(lib().port_max_threads())


# ---------------------------------------------------------------- kernels --
def _qubits_arg(qubits):
    if qubits is None:
        return None, 0, None
    q = np.ascontiguousarray(qubits, dtype=np.int32)
    return _p(q), len(q), q


def apply_gate_kernel(state, gate, nstates, m, qubits=None):
    """gates.py:15-38 (plain and multicontrol)."""
    g = np.ascontiguousarray(np.asarray(gate).reshape(-1), dtype=state.dtype)
    qp, nq, keep = _qubits_arg(qubits)
    getattr(lib(), "port_apply_gate_" + _suffix(state.dtype))(
        _p(state), _p(g), ctypes.c_int64(nstates), ctypes.c_int(m), qp, ctypes.c_int(nq)
    )
    del keep
    return state


_PAULI = {"apply_x": 1, "apply_y": 2, "apply_z": 3, "apply_z_pow": 4}


def apply_pauli_kernel(kind, state, gate, nstates, m, qubits=None):
    """gates.py:41-114."""
    g = np.ascontiguousarray(np.asarray(gate if gate is not None else [0]).reshape(-1), dtype=state.dtype)
    qp, nq, keep = _qubits_arg(qubits)
    getattr(lib(), "port_apply_pauli_" + _suffix(state.dtype))(
        _p(state), _p(g), ctypes.c_int64(nstates), ctypes.c_int(m), qp, ctypes.c_int(nq), ctypes.c_int(_PAULI[kind])
    )
    del keep
    return state


_TWO = {"apply_two_qubit_gate": 0, "apply_swap": 1, "apply_fsim": 2}


def apply_two_kernel(kind, state, gate, nstates, m1, m2, swap_targets, qubits=None):
    """gates.py:117-254."""
    g = np.ascontiguousarray(np.asarray(gate if gate is not None else [0]).reshape(-1), dtype=state.dtype)
    qp, nq, keep = _qubits_arg(qubits)
    getattr(lib(), "port_apply_two_" + _suffix(state.dtype))(
        _p(state),
        _p(g),
        ctypes.c_int64(nstates),
        ctypes.c_int(m1),
        ctypes.c_int(m2),
        ctypes.c_int(int(swap_targets)),
        qp,
        ctypes.c_int(nq),
        ctypes.c_int(_TWO[kind]),
    )
    del keep
    return state


def apply_multi_kernel(state, gate, qubits, nstates, targets):
    """gates.py:265-424; targets = int64 strides in reversed target order."""
    g = np.ascontiguousarray(np.asarray(gate), dtype=state.dtype)
    t = np.ascontiguousarray(targets, dtype=np.int64)
    q = np.ascontiguousarray(qubits, dtype=np.int32)
    getattr(lib(), "port_apply_multi_" + _suffix(state.dtype))(
        _p(state), _p(g), ctypes.c_int64(nstates), _p(t), ctypes.c_int(len(t)), _p(q), ctypes.c_int(len(q))
    )
    return state


def initial_state(state):
    getattr(lib(), "port_initial_state_" + _suffix(state.dtype))(_p(state), ctypes.c_int64(state.size))
    return state


def collapse(state, qubits_kernel, result, nqubits, normalize):
    """ops.py:43-79; returns (state, norm)."""
    q = np.ascontiguousarray(qubits_kernel, dtype=np.int32)
    fn = getattr(lib(), "port_collapse_" + _suffix(state.dtype))
    fn.restype = ctypes.c_double
    norm = fn(_p(state), _p(q), ctypes.c_int(len(q)), ctypes.c_int64(result), ctypes.c_int(nqubits), ctypes.c_int(int(normalize)))
    return state, norm


def measure_frequencies(frequencies, probs, nshots, nqubits, seed, nthreads):
    """ops.py:82-108 (frequencies are added to in place and returned)."""
    del nqubits
    probs = np.ascontiguousarray(probs)
    suf = "f32" if probs.dtype == np.float32 else "f64"
    if probs.dtype not in (np.float32, np.float64):
        probs = probs.astype(np.float64)
    out = np.zeros(len(frequencies), dtype=np.int64)
    getattr(lib(), "port_measure_frequencies_" + suf)(
        _p(out), _p(probs), ctypes.c_int64(len(probs)), ctypes.c_int64(nshots), ctypes.c_uint32(int(seed)), ctypes.c_int(int(nthreads))
    )
    frequencies += out.astype(frequencies.dtype)
    return frequencies


def chain_seeds(seed, nthreads):
    out = np.zeros(nthreads, dtype=np.uint32)
    lib().port_chain_seeds(ctypes.c_uint32(int(seed)), ctypes.c_int(nthreads), _p(out))
    return out


def mt_words(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().port_mt_words(ctypes.c_uint32(seed), _p(out), ctypes.c_int64(n))
    return out


def mt_randint(seed, low, high, n):
    out = np.zeros(n, dtype=np.int64)
    lib().port_mt_randint(ctypes.c_uint32(seed), ctypes.c_int64(low), ctypes.c_int64(high), _p(out), ctypes.c_int64(n))
    return out


def mt_random(seed, n):
    out = np.zeros(n, dtype=np.float64)
    lib().port_mt_random(ctypes.c_uint32(seed), _p(out), ctypes.c_int64(n))
    return out


# ------------------------------------------------------------- dispatcher --
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


def custom_matrix(gate):
    """Kernel-facing encoding (matrices.py:7-70) from textbook matrices."""
    name = gate.__class__.__name__
    p = list(gate.parameters)
    fixed = {
        "H": gm.H, "X": gm.X, "Y": gm.Y, "Z": gm.Z, "S": gm.S, "SDG": gm.SDG, "T": gm.T,
        "TDG": gm.TDG, "SX": gm.SX, "SXDG": gm.SXDG, "CNOT": gm.X, "TOFFOLI": gm.X,
        "CY": gm.Y, "CZ": gm.Z, "CCZ": gm.Z, "CH": gm.H, "CSX": gm.SX, "CSXDG": gm.SXDG,
        "SWAP": gm.SWAP,
    }
    if name in fixed:
        return fixed[name]
    if name in ("RX", "CRX"):
        return gm.RX(p[0])
    if name in ("RY", "CRY"):
        return gm.RY(p[0])
    if name in ("RZ", "CRZ"):
        return gm.RZ(p[0])
    if name in ("U1", "CU1"):
        return np.exp(1j * p[0])
    if name in ("U2", "CU2"):
        return gm.U2(*p)
    if name in ("U3", "CU3"):
        return gm.U3(*p)
    if name == "DEUTSCH":
        return 1j * gm.RX(2 * p[0])
    if name == "fSim":
        c, s = np.cos(p[0]), -1j * np.sin(p[0])
        return np.array([c, s, s, c, np.exp(-1j * p[1])])
    if name == "GeneralizedfSim":
        u = np.asarray(p[0])
        return np.array([u[0, 0], u[0, 1], u[1, 0], u[1, 1], np.exp(-1j * p[1])])
    if name == "Unitary":
        return np.asarray(p[0])
    raise NotImplementedError(name)


class PortBackend:
    """NumbaBackend's dispatch (cpu.py:369-635) over the C port, on numpy."""

    def __init__(self, nthreads=None):
        self.nthreads = nthreads or len(os.sched_getaffinity(0))
        set_threads(self.nthreads)

    @staticmethod
    def _create_qubits_tensor(gate, nqubits):
        qubits = [nqubits - q - 1 for q in gate.control_qubits]
        qubits.extend(nqubits - q - 1 for q in gate.target_qubits)
        return np.array(sorted(qubits), dtype=np.int32)

    def apply_gate(self, gate, state, nqubits, inverse=False):
        if state.ndim == 2:
            return self._apply_gate_density_matrix(gate, state, nqubits, inverse)
        return self._apply_gate(gate, state, nqubits)

    def _apply_gate(self, gate, state, nqubits):
        if gate.name == "fanout":
            control = gate.control_qubits[0]
            for target in gate.target_qubits:
                qubits = np.array(sorted([nqubits - control - 1, nqubits - target - 1]), dtype=np.int32)
                state = self._one_qubit_base(state, nqubits, target, "apply_x", None, qubits)
            return state
        matrix = custom_matrix(gate)
        qubits = self._create_qubits_tensor(gate, nqubits)
        targets = gate.target_qubits
        name = gate.__class__.__name__
        if len(targets) == 1:
            return self._one_qubit_base(state, nqubits, *targets, GATE_OPS.get(name, "apply_gate"), matrix, qubits)
        if len(targets) == 2:
            return self._two_qubit_base(state, nqubits, *targets, GATE_OPS.get(name, "apply_two_qubit_gate"), matrix, qubits)
        return self._multi_qubit_base(state, nqubits, targets, matrix, qubits)

    def _apply_gate_density_matrix(self, gate, state, nqubits, inverse=False):
        name = gate.__class__.__name__
        matrix = custom_matrix(gate)
        if inverse:
            matrix = np.linalg.inv(np.asarray(matrix))
        qubits = self._create_qubits_tensor(gate, nqubits)
        qubits_dm = qubits + nqubits
        targets = gate.target_qubits
        targets_dm = tuple(q + nqubits for q in targets)
        shape = state.shape
        flat = state.reshape(-1)
        if name in ("Y", "CY"):
            flat = self._one_qubit_base(flat, 2 * nqubits, *targets, "apply_y", matrix, qubits_dm)
            flat = self._one_qubit_base(flat, 2 * nqubits, *targets_dm, "apply_gate", np.conj(matrix), qubits)
            return flat.reshape(shape)
        if len(targets) == 1:
            op = GATE_OPS.get(name, "apply_gate")
            flat = self._one_qubit_base(flat, 2 * nqubits, *targets, op, matrix, qubits_dm)
            flat = self._one_qubit_base(flat, 2 * nqubits, *targets_dm, op, np.conj(matrix), qubits)
        elif len(targets) == 2:
            op = GATE_OPS.get(name, "apply_two_qubit_gate")
            flat = self._two_qubit_base(flat, 2 * nqubits, *targets, op, matrix, qubits_dm)
            flat = self._two_qubit_base(flat, 2 * nqubits, *targets_dm, op, np.conj(matrix), qubits)
        else:
            flat = self._multi_qubit_base(flat, 2 * nqubits, targets, matrix, qubits_dm)
            flat = self._multi_qubit_base(flat, 2 * nqubits, targets_dm, np.conj(matrix), qubits)
        return flat.reshape(shape)

    @staticmethod
    def _one_qubit_base(state, nqubits, target, kernel, gate, qubits):
        ncontrols = len(qubits) - 1 if qubits is not None else 0
        m = nqubits - target - 1
        nstates = 1 << (nqubits - ncontrols - 1)
        q = qubits if ncontrols else None
        if kernel == "apply_gate":
            return apply_gate_kernel(state, gate, nstates, m, q)
        return apply_pauli_kernel(kernel, state, gate, nstates, m, q)

    @staticmethod
    def _two_qubit_base(state, nqubits, target1, target2, kernel, gate, qubits):
        ncontrols = len(qubits) - 2 if qubits is not None else 0
        if target1 > target2:
            swap_targets = True
            m1, m2 = nqubits - target1 - 1, nqubits - target2 - 1
        else:
            swap_targets = False
            m1, m2 = nqubits - target2 - 1, nqubits - target1 - 1
        nstates = 1 << (nqubits - 2 - ncontrols)
        q = qubits if ncontrols else None
        return apply_two_kernel(kernel, state, gate, nstates, m1, m2, swap_targets, q)

    @staticmethod
    def _multi_qubit_base(state, nqubits, targets, gate, qubits):
        if qubits is None:
            qubits = np.array(sorted(nqubits - q - 1 for q in targets), dtype=np.int32)
        nstates = 1 << (nqubits - len(qubits))
        strides = [1 << (nqubits - t - 1) for t in list(targets)[::-1]]
        return apply_multi_kernel(state, gate, qubits, nstates, strides)

    def zero_state(self, nqubits, density_matrix=False, dtype=np.complex128):
        n = 2**nqubits
        state = np.empty(n * n if density_matrix else n, dtype=dtype)
        initial_state(state)
        return state.reshape(n, n) if density_matrix else state

    def collapse_state(self, state, qubits, shot, nqubits, normalize=True):
        qk = [nqubits - q - 1 for q in reversed(list(qubits))]
        state, _ = collapse(state, qk, int(shot), nqubits, normalize)
        return state

    def execute_circuit(self, circuit, initial_state=None):
        n = circuit.nqubits
        state = self.zero_state(n) if initial_state is None else np.array(initial_state, copy=True)
        for gate in circuit.queue:
            if gate.__class__.__name__ == "FusedGate":
                for g in gate.gates:
                    state = self.apply_gate(g, state, n)
            else:
                state = self.apply_gate(gate, state, n)
        return state
