"""Stabilizer-tableau operations on the GPU: the engine module of a Clifford
backend, in the role of src/qibojit/backends/clifford_operations_gpu.py.

The module-level names, their arguments and their results follow the
reference module: ``H(symplectic_matrix, q, nqubits)`` ...
``CY(symplectic_matrix, control_q, target_q, nqubits)`` update the tableau
and return it, ``_rowsum``, ``_random_outcome``, ``_determined_outcome``,
``_init_state_for_measurements``, ``_clifford_pre_execution_reshape`` /
``_clifford_post_execution_reshape``, ``_packbits`` / ``_unpackbits`` and
``_pack_for_measurements`` / ``_unpack_for_measurements``.

Between the pre- and post-execution reshapes the state is a device tableau:
an int32 tensor of shape (2n+1, ws), column c of the reference's
(2n+1) x (2n+1) 0/1 matrix bit-packed over the rows (row j = bit j % 32 of
word j / 32; ws = ``column_words(n)``).  The reference's GPU engine packs the
matrix over rows too (``_packbits(state, axis=0)``,
clifford_operations_gpu.py:694-695) but keeps one byte per column of 8 rows;
here a column is a contiguous bit vector, so a gate reads and writes three
(five) contiguous vectors and ``apply_circuit`` runs a whole gate list in one
launch.  Every computation is a kernel of libqjhip.so (no host fallback).
"""

import ctypes

import numpy as np
import torch

from qibojit_amd import _native

KINDS = {"H": 0, "S": 1, "Z": 2, "X": 3, "Y": 4, "SX": 5, "SDG": 6, "SXDG": 7, "RY_pi": 8, "RY_3pi_2": 9,
         "CNOT": 10, "CZ": 11, "SWAP": 12, "iSWAP": 13, "CY": 14}
TWO_QUBIT = ("CNOT", "CZ", "SWAP", "iSWAP", "CY")
GROUP_WORDS = 32  # QJ_CLIFF_GROUP_WORDS


def _stream(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("the Clifford engine needs a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def column_words(nqubits):
    """Words per tableau column for ``nqubits`` qubits."""
    ws = _native.lib().qj_clifford_column_words(int(nqubits))
    if ws < 0:
        raise ValueError(f"nqubits {nqubits} out of range for a tableau")
    return int(ws)


def _nqubits_of_dense(state):
    rows = state.shape[0]
    if state.ndim != 2 or state.shape[1] != rows or rows % 2 != 1 or rows < 3:
        raise ValueError(f"expected a (2n+1) x (2n+1) symplectic matrix, got shape {tuple(state.shape)}")
    return (rows - 1) // 2


def _check_tableau(tab, nqubits):
    if not isinstance(tab, torch.Tensor) or tab.device.type != "cuda" or tab.dtype != torch.int32:
        raise TypeError("expected the device tableau (int32 CUDA tensor) of _clifford_pre_execution_reshape")
    shape = (2 * int(nqubits) + 1, column_words(nqubits))
    if tuple(tab.shape) != shape or not tab.is_contiguous():
        raise ValueError(f"tableau of shape {tuple(tab.shape)} does not match {nqubits} qubits (expected {shape})")


def _as_device_bytes(array):
    if isinstance(array, torch.Tensor):
        t = array.to(device=_device(), dtype=torch.uint8)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(array), dtype=np.uint8)).to(_device())
    return t.contiguous()


# ---- layout ----

def _clifford_pre_execution_reshape(state):
    """(2n+1) x (2n+1) 0/1 matrix (host or device) -> device tableau."""
    dense = _as_device_bytes(state)
    n = _nqubits_of_dense(dense)
    tab = torch.empty((2 * n + 1, column_words(n)), dtype=torch.int32, device=dense.device)
    _native.call("qj_clifford_pack", _ptr(dense), n, _ptr(tab), _stream(tab))
    return tab


def _clifford_post_execution_reshape(state, nqubits):
    """Device tableau -> (2n+1) x (2n+1) uint8 device matrix."""
    _check_tableau(state, nqubits)
    dim = 2 * int(nqubits) + 1
    dense = torch.empty((dim, dim), dtype=torch.uint8, device=state.device)
    _native.call("qj_clifford_unpack", _ptr(state), int(nqubits), _ptr(dense), _stream(state))
    return dense


def zero_state(nqubits):
    """Device tableau of |0...0> (destabilizer k = X_k, stabilizer k = Z_k)."""
    n = int(nqubits)
    dense = torch.zeros((2 * n + 1, 2 * n + 1), dtype=torch.uint8, device=_device())
    idx = torch.arange(2 * n, device=dense.device)
    dense[idx, idx] = 1
    return _clifford_pre_execution_reshape(dense)


# ---- gates (clifford_operations_gpu.py:49-594) ----

def _apply(kind, tab, q0, q1, nqubits):
    _check_tableau(tab, nqubits)
    _native.call("qj_clifford_apply_gate", _ptr(tab), int(nqubits), kind, int(q0), int(q1), _stream(tab))
    return tab


def _one_qubit(name):
    kind = KINDS[name]

    def gate(symplectic_matrix, q, nqubits):
        return _apply(kind, symplectic_matrix, q, -1, nqubits)

    gate.__name__ = name
    gate.__doc__ = f"{name} on qubit q of every row of the tableau."
    return gate


def _two_qubit(name):
    kind = KINDS[name]

    def gate(symplectic_matrix, control_q, target_q, nqubits):
        return _apply(kind, symplectic_matrix, control_q, target_q, nqubits)

    gate.__name__ = name
    gate.__doc__ = f"{name} on (control_q, target_q) of every row of the tableau."
    return gate


H = _one_qubit("H")
S = _one_qubit("S")
Z = _one_qubit("Z")
X = _one_qubit("X")
Y = _one_qubit("Y")
SX = _one_qubit("SX")
SDG = _one_qubit("SDG")
SXDG = _one_qubit("SXDG")
RY_pi = _one_qubit("RY_pi")
RY_3pi_2 = _one_qubit("RY_3pi_2")
CNOT = _two_qubit("CNOT")
CZ = _two_qubit("CZ")
SWAP = _two_qubit("SWAP")
iSWAP = _two_qubit("iSWAP")
CY = _two_qubit("CY")


def encode_gates(gates):
    """(name, q) / (name, control, target) sequence -> int32 (ngates, 3)."""
    out = np.empty((len(gates), 3), dtype=np.int32)
    for j, g in enumerate(gates):
        name, qubits = g[0], g[1:]
        if name not in KINDS:
            raise ValueError(f"{name} is not a Clifford gate of the tableau engine")
        want = 2 if name in TWO_QUBIT else 1
        if len(qubits) != want:
            raise ValueError(f"{name} takes {want} qubit(s), got {len(qubits)}")
        out[j, 0] = KINDS[name]
        out[j, 1] = qubits[0]
        out[j, 2] = qubits[1] if want == 2 else -1
    return out


def apply_circuit(symplectic_matrix, gates, nqubits):
    """Applies a gate list in one launch; equal to applying the gates one by
    one.  ``gates`` is a sequence of (name, q) / (name, control, target) or
    an int32 array of (kind, q0, q1) rows."""
    _check_tableau(symplectic_matrix, nqubits)
    arr = gates if isinstance(gates, np.ndarray) and gates.dtype == np.int32 else encode_gates(gates)
    arr = np.ascontiguousarray(arr.reshape(-1, 3))
    ptr = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    _native.call("qj_clifford_apply_circuit", _ptr(symplectic_matrix), int(nqubits), ptr, arr.shape[0],
                 _stream(symplectic_matrix))
    return symplectic_matrix


class CliffordSchedule:
    """A gate list ordered into groups once (``compile_circuit``), held in
    device memory, applied with ``apply_schedule`` as often as needed."""

    def __init__(self, desc, ngroups, ngates, nqubits):
        self.desc = desc
        self.ngroups = ngroups
        self.ngates = ngates
        self.nqubits = nqubits


def compile_circuit(gates, nqubits):
    """Orders a gate list (as for ``apply_circuit``) into its device schedule."""
    arr = gates if isinstance(gates, np.ndarray) and gates.dtype == np.int32 else encode_gates(gates)
    arr = np.ascontiguousarray(arr.reshape(-1, 3))
    ptr = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    ngroups = ctypes.c_int64(0)
    _native.call("qj_clifford_schedule", ptr, arr.shape[0], int(nqubits), None, 0, ctypes.byref(ngroups))
    words = np.empty(max(1, ngroups.value) * GROUP_WORDS, dtype=np.uint32)
    _native.call("qj_clifford_schedule", ptr, arr.shape[0], int(nqubits), words.ctypes.data_as(ctypes.c_void_p),
                 ngroups.value, ctypes.byref(ngroups))
    desc = torch.from_numpy(words.view(np.int32)).to(_device())
    return CliffordSchedule(desc, ngroups.value, arr.shape[0], int(nqubits))


def apply_schedule(symplectic_matrix, schedule, nqubits):
    """Applies a ``CliffordSchedule`` in one launch."""
    _check_tableau(symplectic_matrix, nqubits)
    if schedule.nqubits != int(nqubits):
        raise ValueError(f"schedule compiled for {schedule.nqubits} qubits, tableau has {nqubits}")
    if schedule.desc.device != symplectic_matrix.device:
        raise ValueError("schedule and tableau are on different devices")
    _native.call("qj_clifford_apply_schedule", _ptr(symplectic_matrix), int(nqubits), _ptr(schedule.desc),
                 schedule.ngroups, _stream(symplectic_matrix))
    return symplectic_matrix


# ---- rowsum and measurement ----

def _rowsum(symplectic_matrix, h, i, nqubits, determined=False):
    """Rows h[j] <- P_i[j] P_h[j] (clifford_operations_cpu.py:337-385); in the
    determined mode every product is XORed into row h[0]."""
    _check_tableau(symplectic_matrix, nqubits)
    h = np.ascontiguousarray(np.asarray(h, dtype=np.int64).reshape(-1))
    i = np.ascontiguousarray(np.asarray(i, dtype=np.int64).reshape(-1))
    if h.shape != i.shape:
        raise ValueError("h and i must have the same length")
    _native.call("qj_clifford_rowsum", _ptr(symplectic_matrix), int(nqubits), h.ctypes.data_as(ctypes.c_void_p),
                 i.ctypes.data_as(ctypes.c_void_p), h.shape[0], int(bool(determined)), _stream(symplectic_matrix))
    return symplectic_matrix


def first_stabilizer(state, q, nqubits):
    """First stabilizer (0-based) whose x bit on q is 1, or -1 (synchronises)."""
    _check_tableau(state, nqubits)
    out = torch.empty(1, dtype=torch.int32, device=state.device)
    _native.call("qj_clifford_first_stabilizer", _ptr(state), int(nqubits), int(q), _ptr(out), _stream(state))
    return int(out.item())


def _random_outcome(state, p, q, nqubits):
    """Random outcome of qubit q (clifford_operations_gpu.py:606-625); p[0] is
    the first anticommuting stabilizer.  The bit is drawn from NumPy's global
    generator."""
    _check_tableau(state, nqubits)
    p0 = int(np.asarray(p).reshape(-1)[0])
    outcome = np.random.randint(2, size=None, dtype=np.uint)
    _native.call("qj_clifford_random_outcome", _ptr(state), int(nqubits), p0, int(q), int(outcome), _stream(state))
    return state, outcome


def _determined_outcome(state, q, nqubits, exact=False):
    """Determined outcome of qubit q (clifford_operations_gpu.py:628-640):
    the XOR of the selected stabilizers' signs.  ``exact=True`` (no reference
    counterpart) accumulates the phases of their product instead, which is
    always the state's certain outcome; the XOR rule can differ from it."""
    _check_tableau(state, nqubits)
    out = torch.empty(1, dtype=torch.int32, device=state.device)
    name = "qj_clifford_determined_outcome_exact" if exact else "qj_clifford_determined_outcome"
    _native.call(name, _ptr(state), int(nqubits), int(q), _ptr(out), _stream(state))
    return state, np.uint8(out.item())


def _init_state_for_measurements(state, nqubits, collapse):
    """The tableau measured in place when collapsing, a copy otherwise."""
    _check_tableau(state, nqubits)
    return state if collapse else state.clone()


def M(state, qubits, nqubits, collapse=False, exact=False):
    """Measures ``qubits`` one after another (the measurement loop of qibo's
    Clifford backend, which calls this module's outcome functions);
    ``exact`` selects the determined-outcome rule of ``_determined_outcome``.

    The loop runs on the device in one call (``qj_clifford_measure``) and
    reads back once.  The random bits come from NumPy's global generator as
    in the loop of ``_random_outcome`` calls: len(qubits) bits are drawn up
    front, and the generator is then rewound and advanced by the number of
    random outcomes actually taken (a vector draw of randint(2) consumes the
    stream as the same number of scalar draws)."""
    state = _init_state_for_measurements(state, nqubits, collapse)
    qs = np.asarray([int(q) for q in qubits], dtype=np.int32)
    m = len(qs)
    if m == 0:
        return []
    saved = np.random.get_state()
    draws = np.random.randint(2, size=m, dtype=np.uint).astype(np.uint8)
    dev = state.device
    draws_d = torch.from_numpy(draws).to(dev)
    out = torch.empty(m + 1, dtype=torch.int32, device=dev)
    _native.call("qj_clifford_measure", _ptr(state), int(nqubits), qs.ctypes.data_as(ctypes.c_void_p), m,
                 _ptr(draws_d), int(bool(exact)), _ptr(out), _ptr(out[m:]), _stream(state))
    host = out.cpu().numpy()
    np.random.set_state(saved)
    np.random.randint(2, size=int(host[m]), dtype=np.uint)
    return [int(v) for v in host[:m]]


def measure_with_outcomes(state, qubits, nqubits, ones, exact=False):
    """The measurement loop of ``M`` on ``state`` (collapsed in place) with
    the random outcomes chosen instead of drawn (post-selection): the t-th
    random outcome (t = 0, 1, ...) is 1 iff t is in ``ones``.  Returns the
    sample (uint8 array) and the positions in ``qubits`` of the random
    outcomes."""
    _check_tableau(state, nqubits)
    sample, random_at = [], []
    for pos, q in enumerate(qubits):
        p = first_stabilizer(state, q, nqubits)
        if p >= 0:
            bit = int(len(random_at) in ones)
            _native.call("qj_clifford_random_outcome", _ptr(state), int(nqubits), p, int(q), bit, _stream(state))
            random_at.append(pos)
            sample.append(bit)
        else:
            _, outcome = _determined_outcome(state, q, nqubits, exact)
            sample.append(int(outcome))
    return np.array(sample, dtype=np.uint8), random_at


def measurement_forms(state, qubits, nqubits, exact=False):
    """Affine forms over GF(2) of the bits ``M(state, qubits, nqubits, exact=exact)``
    would return: a (len(qubits), k + 1) uint8 array F and k, the number of
    random outcomes, such that the sample for random outcomes u_1..u_k is
    F[:, 0] ^ (F[:, 1:] . u mod 2).  One symbolic measurement pass (the
    state is left unchanged).

    For a fixed state, which measurements are random, and every x/z bit on
    the way, do not depend on the outcomes (``first_stabilizer`` and the
    rowsum's x/z updates read x/z only).  The signs do, affinely, so they are
    carried as forms: sign column j of the device buffer holds every row's
    coefficient of term j."""
    _check_tableau(state, nqubits)
    qubits = [int(q) for q in qubits]
    m, n = len(qubits), int(nqubits)
    tab = state.clone()
    signs = torch.empty(((m + 1), column_words(n)), dtype=torch.int32, device=state.device)
    forms = torch.zeros((m, m + 1), dtype=torch.uint8, device=state.device)
    stream = _stream(state)
    _native.call("qj_clifford_sym_init", _ptr(tab), n, _ptr(signs), m, stream)
    k = 0
    for i, q in enumerate(qubits):
        p = first_stabilizer(tab, q, n)
        if p >= 0:
            k += 1
            _native.call("qj_clifford_sym_random", _ptr(tab), n, _ptr(signs), m, p, q, k, stream)
            forms[i, k] = 1
        else:
            name = "qj_clifford_sym_determined_exact" if exact else "qj_clifford_sym_determined"
            _native.call(name, _ptr(tab), n, _ptr(signs), m, q, _ptr(forms[i]), stream)
    return forms[:, : k + 1].cpu().numpy(), k


def sample_shots(state, qubits, nqubits, nshots, exact=False):
    """``nshots`` measurements of ``qubits`` on copies of ``state`` (left
    unchanged): a (nshots, len(qubits)) uint8 device tensor equal to ``nshots``
    calls of ``M(state, qubits, nqubits, exact=exact)``, with the same draws from NumPy's
    generator (one per random outcome, shot after shot).  The affine forms
    of the measured bits come from one symbolic pass
    (``measurement_forms``); shot s is F[:, 0] ^ (F[:, 1:] . u_s), evaluated
    by ``qj_gf2_samples`` for all shots at once."""
    forms, k = measurement_forms(state, qubits, nqubits, exact)
    m, nshots = forms.shape[0], int(nshots)
    kw = max(1, (k + 63) // 64)
    dt = np.zeros((m, kw * 8), dtype=np.uint8)
    if k:
        packed = np.packbits(forms[:, 1:], axis=1, bitorder="little")
        dt[:, : packed.shape[1]] = packed
    draws = np.random.randint(2, size=(nshots, k), dtype=np.uint).astype(np.uint8)
    u = np.zeros((nshots, kw * 8), dtype=np.uint8)
    if k:
        packed = np.packbits(draws, axis=1, bitorder="little")
        u[:, : packed.shape[1]] = packed
    dev = state.device
    u_d = torch.from_numpy(u.view(np.int64)).to(dev)
    dt_d = torch.from_numpy(dt.view(np.int64)).to(dev)
    s0_d = torch.from_numpy(np.ascontiguousarray(forms[:, 0])).to(dev)
    out = torch.empty((nshots, m), dtype=torch.uint8, device=dev)
    _native.call("qj_gf2_samples", _ptr(u_d), _ptr(dt_d), _ptr(s0_d), nshots, m, kw, _ptr(out), _stream(state))
    return out


# ---- numpy-format bit packing (clifford_operations_cpu.py:250-335) ----

def _packed_size(n):
    return (int(n) + 7) // 8


def _packbits(array, axis):
    """numpy.packbits(array, axis) of a 2-D 0/1 matrix, on the device."""
    a = _as_device_bytes(array)
    if a.ndim != 2 or axis not in (0, 1, -1, -2):
        raise ValueError("_packbits takes a 2-D matrix and axis 0 or 1")
    if axis in (0, -2):
        a = a.t().contiguous()
    rows, cols = a.shape
    out = torch.empty((rows, _packed_size(cols)), dtype=torch.uint8, device=a.device)
    _native.call("qj_packbits_rows", _ptr(a), rows, cols, _ptr(out), _stream(a))
    return out.t().contiguous() if axis in (0, -2) else out


def _unpackbits(array, axis, count):
    """numpy.unpackbits(array, axis, count=count) of a 2-D byte matrix."""
    a = _as_device_bytes(array)
    if a.ndim != 2 or axis not in (0, 1, -1, -2):
        raise ValueError("_unpackbits takes a 2-D matrix and axis 0 or 1")
    if axis in (0, -2):
        a = a.t().contiguous()
    rows, pcols = a.shape
    out = torch.empty((rows, int(count)), dtype=torch.uint8, device=a.device)
    _native.call("qj_unpackbits_rows", _ptr(a), rows, pcols, int(count), _ptr(out), _stream(a))
    return out.t().contiguous() if axis in (0, -2) else out


def _pack_for_measurements(state, nqubits):
    """Dense matrix -> [packbits(x), packbits(z), r] per row."""
    a = _as_device_bytes(state)
    n = int(nqubits)
    return torch.hstack((_packbits(a[:, :n], axis=1), _packbits(a[:, n:-1], axis=1), a[:, -1:]))


def _unpack_for_measurements(state, nqubits):
    """Inverse of _pack_for_measurements."""
    a = _as_device_bytes(state)
    n, pn = int(nqubits), _packed_size(nqubits)
    return torch.hstack((_unpackbits(a[:, :pn], axis=1, count=n), _unpackbits(a[:, pn:-1], axis=1, count=n),
                         a[:, -1:]))
