"""HipBackend: qibo's state-vector backend interface on MI355X (gfx950).

The host side mirrors qibojit's NumbaBackend/CupyBackend dispatch
(src/qibojit/backends/cpu.py:369-635, gpu.py:584-1076): a gate becomes a
kernel family (GATE_OPS), a kernel-facing matrix (KernelMatrices), a sorted
"qubits tensor" of bit positions and a target count, and then one call into
``libqjhip.so``.  Unlike CupyBackend, gate matrices and qubit tensors stay on
the host (they travel as kernel arguments or through the library's staging
ring) and no call synchronises the host per gate.

States are torch tensors on the current HIP device: torch provides device
memory and the stream, every state-vector computation is a kernel of the
library.  There is no CPU fallback; without the library or a device the
constructor raises ImportError, as CupyBackend does without CUDA
(gpu.py:85-91).
"""

import ctypes
import math
import os
from collections import Counter

import numpy as np
import torch

from qibojit_amd import _native
from qibojit_amd._native import ONE_QUBIT_KINDS, TWO_QUBIT_KINDS, call, i32_array
from qibojit_amd._qibo import (
    SHOT_METROPOLIS_THRESHOLD,
    Backend,
    CircuitResult,
    FusedGate,
    NumpyMatrices,
    QuantumState,
    raise_error,
)
from qibojit_amd import tiles
from qibojit_amd.backends.matrices import KernelMatrices
from qibojit_amd.fusion import GATE_OPS, fused_matrix


_TORCH_DTYPES = {
    "complex64": torch.complex64,
    "complex128": torch.complex128,
    "float32": torch.float32,
    "float64": torch.float64,
    "int8": torch.int8,
    "int32": torch.int32,
    "int64": torch.int64,
}
_NATIVE_CODES = {
    torch.complex64: _native.QJ_COMPLEX64,
    torch.complex128: _native.QJ_COMPLEX128,
    torch.float32: _native.QJ_FLOAT32,
    torch.float64: _native.QJ_FLOAT64,
}
_REAL_OF = {torch.complex64: torch.float32, torch.complex128: torch.float64}


def _dtype_name(dtype):
    if isinstance(dtype, torch.dtype):
        return str(dtype).replace("torch.", "")
    return str(np.dtype(dtype))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


# The current stream of a device as a raw pointer without building a Stream
# object (torch.cuda.current_stream costs a few microseconds per launch)
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _device_empty(numel, tdtype, device):
    """Uninitialised device tensor; an allocation the device cannot hold
    raises MemoryError (numpy's error for the same case on NumbaBackend),
    checked against the device's capacity before any allocation is tried."""
    nbytes = int(numel) * (torch.empty((), dtype=tdtype).element_size())
    capacity = torch.cuda.get_device_properties(device).total_memory
    if nbytes > capacity:
        raise MemoryError(f"{nbytes / 2**30:.1f} GiB do not fit the {capacity / 2**30:.1f} GiB of {device}.")
    try:
        return torch.empty(int(numel), dtype=tdtype, device=device)
    except torch.cuda.OutOfMemoryError as exc:
        raise MemoryError(f"device allocation of {nbytes / 2**30:.1f} GiB failed: {exc}") from exc


class HipBackend(Backend):
    """qibo Backend whose state-vector operations run as gfx950 HIP kernels."""

    MAX_NUM_TARGETS = 12

    def __init__(self):
        super().__init__()
        from qibojit_amd import __version__  # pylint: disable=import-outside-toplevel

        try:
            _native.lib()
        except OSError as exc:
            raise ImportError(f"Could not load libqjhip.so: {exc}") from exc
        if not torch.cuda.is_available() or _native.device_count() == 0:
            raise ImportError("Could not detect HIP compatible devices.")
        self.name = "qibojit"
        self.platform = "hip"
        self.engine = torch
        self.np = np
        self.dtype = "complex128"
        self.device = "/GPU:0"
        self._torch_device = torch.device("cuda", 0)
        self.matrices = NumpyMatrices(self.dtype)
        self.custom_matrices = KernelMatrices(self.dtype)
        self.tensor_types = (np.ndarray, torch.Tensor)
        self.supports_multigpu = False
        self.oom_error = (MemoryError, torch.cuda.OutOfMemoryError)
        self.versions.update(
            {"qibojit_amd": __version__, "torch": torch.__version__, "hip": str(torch.version.hip)}
        )
        self.measure_frequencies_op = self.measure_frequencies
        # Metropolis chain count, as NumbaBackend (cpu.py:86-89).
        self.nthreads = len(os.sched_getaffinity(0))
        # execute_circuit runs unitary segments through the fused tile executor
        self.fusion = True
        # uncontrolled 4/5-target complex64 blocks run on FP32 matrix cores
        self.mfma = True
        self._custom_matrices_c128 = KernelMatrices("complex128")

    # ------------------------------------------------------------ config --
    def set_device(self, device):
        if "GPU" not in device:
            raise_error(ValueError, f"Device {device} is not available for {self} backend.")
        idx = int(device.split(":")[-1]) if ":" in device else 0
        if idx >= torch.cuda.device_count():
            raise_error(ValueError, f"Device {device} is not available for {self} backend.")
        self.device = device
        self._torch_device = torch.device("cuda", idx)
        torch.cuda.set_device(self._torch_device)
        call("qj_set_device", idx)

    def set_dtype(self, dtype):
        dtype = _dtype_name(dtype)
        if dtype != self.dtype:
            super().set_dtype(dtype)
            self.matrices = NumpyMatrices(self.dtype)
            self.custom_matrices = KernelMatrices(self.dtype)

    def set_seed(self, seed):
        super().set_seed(seed)

    def set_threads(self, nthreads):
        self.nthreads = int(nthreads)

    # ------------------------------------------------------------ arrays --
    def _stream(self):
        if _RAW_STREAM is not None:
            return ctypes.c_void_p(_RAW_STREAM(self._torch_device.index))
        return ctypes.c_void_p(torch.cuda.current_stream(self._torch_device).cuda_stream)

    def cast(self, array, dtype=None, copy=False):
        if dtype is None:
            dtype = self.dtype
        tdt = _TORCH_DTYPES[_dtype_name(dtype)]
        try:
            if isinstance(array, torch.Tensor):
                out = array.to(device=self._torch_device, dtype=tdt)
                if copy and out.data_ptr() == array.data_ptr():
                    out = out.clone()
                return out.contiguous()
            arr = np.asarray(array)
            if arr.dtype.kind == "c" and tdt in (torch.float32, torch.float64):
                arr = arr.real
            arr = np.ascontiguousarray(arr, dtype=_dtype_name(tdt))
            return torch.from_numpy(arr).to(self._torch_device)
        except torch.cuda.OutOfMemoryError as exc:
            raise MemoryError(f"cast: device allocation failed: {exc}") from exc

    def to_numpy(self, array):
        if isinstance(array, torch.Tensor):
            return array.detach().cpu().numpy()
        if isinstance(array, QuantumState):
            return self.to_numpy(array.state())
        return np.asarray(array)

    def is_sparse(self, array):  # pylint: disable=unused-argument
        return False

    def nqubits_of(self, state):
        n = state.shape[0]
        return int(round(math.log2(n)))

    def _as_state(self, state):
        """Contiguous complex device tensor holding ``state`` (no copy if it is one)."""
        if (
            isinstance(state, torch.Tensor)
            and state.device == self._torch_device
            and state.dtype in (torch.complex64, torch.complex128)
            and state.is_contiguous()
        ):
            return state
        dtype = _dtype_name(state.dtype) if hasattr(state, "dtype") else self.dtype
        if dtype not in ("complex64", "complex128"):
            dtype = self.dtype
        return self.cast(state, dtype=dtype)

    @staticmethod
    def _host_matrix(matrix, tdtype):
        """Kernel-facing matrix as a contiguous host array of the state's precision."""
        if isinstance(matrix, torch.Tensor):
            matrix = matrix.detach().cpu().numpy()
        npdt = np.complex64 if tdtype == torch.complex64 else np.complex128
        return np.ascontiguousarray(np.asarray(matrix), dtype=npdt)

    # -------------------------------------------------------- state prep --
    def zero_state(self, nqubits, density_matrix=False, dtype=None):
        if dtype is None:
            dtype = self.dtype
        tdt = _TORCH_DTYPES[_dtype_name(dtype)]
        n = 1 << nqubits
        numel = n * n if density_matrix else n
        state = _device_empty(numel, tdt, self._torch_device)
        call("qj_initial_state", _ptr(state), _NATIVE_CODES[tdt], numel, self._stream())
        return state.reshape(n, n) if density_matrix else state

    def plus_state(self, nqubits, density_matrix=False, dtype=None):
        """|+...+> (all amplitudes 2^(-n/2)), or its density matrix (all 2^-n)."""
        if dtype is None:
            dtype = self.dtype
        tdt = _TORCH_DTYPES[_dtype_name(dtype)]
        n = 1 << nqubits
        numel = n * n if density_matrix else n
        state = _device_empty(numel, tdt, self._torch_device)
        value = 1.0 / n if density_matrix else 1.0 / math.sqrt(n)
        call("qj_fill_state", _ptr(state), _NATIVE_CODES[tdt], numel, value, 0.0, self._stream())
        return state.reshape(n, n) if density_matrix else state

    # ---------------------------------------------------- gate dispatch --
    def apply_gate(self, gate, state, nqubits, inverse=False):
        density_matrix = len(tuple(state.shape)) == 2
        if density_matrix:
            return self._apply_gate_density_matrix(gate, state, nqubits, inverse)
        return self._apply_gate(gate, state, nqubits)

    def apply_gate_density_matrix(self, gate, state, nqubits, inverse=False):
        return self._apply_gate_density_matrix(gate, state, nqubits, inverse)

    def apply_gate_half_density_matrix(self, gate, state, nqubits):
        """rho -> U rho: the row pass of the density-matrix update only."""
        if len(gate.control_qubits) > 0:
            raise_error(
                NotImplementedError,
                "Gate density matrix half call is not implemented for ``controlled_by`` gates.",
            )
        state = self._as_state(state)
        shape = state.shape
        matrix = self._as_custom_matrix(gate)
        qubits = self._create_qubits_tensor(gate, nqubits)
        targets = gate.target_qubits
        flat = state.reshape(-1)
        name = gate.__class__.__name__
        if len(targets) == 1:
            op = GATE_OPS.get(name, "apply_gate")
            flat = self._one_qubit_base(flat, 2 * nqubits, *targets, op, matrix, qubits + nqubits)
        elif len(targets) == 2:
            op = GATE_OPS.get(name, "apply_two_qubit_gate")
            flat = self._two_qubit_base(flat, 2 * nqubits, *targets, op, matrix, qubits + nqubits)
        else:
            flat = self._multi_qubit_base(flat, 2 * nqubits, targets, matrix, qubits + nqubits)
        return flat.reshape(shape)

    def apply_channel(self, channel, state, nqubits):
        density_matrix = len(tuple(state.shape)) == 2
        if density_matrix:
            return self._apply_channel_density_matrix(channel, state, nqubits)
        raise_error(NotImplementedError, "Channels on state vectors are sampled by qibo, not by this backend.")

    def _apply_channel_density_matrix(self, channel, state, nqubits):
        """sum_i p_i U_i rho U_i^+ + (1 - sum p_i) rho (cpu.py:400-415)."""
        state = self._as_state(state)
        if not channel._all_unitary_operators:  # pylint: disable=protected-access
            state_copy = state.clone()
        new_state = torch.empty_like(state)
        self._axpby(state, new_state, 1 - channel.coefficient_sum, 0.0)
        for coeff, gate in zip(channel.coefficients, channel.gates):
            state = self.apply_gate(gate, state, nqubits)
            self._axpby(state, new_state, coeff, 1.0)
            if not channel._all_unitary_operators:  # pylint: disable=protected-access
                state = state_copy.clone()
            else:
                state = self.apply_gate(gate, state, nqubits, inverse=True)
        return new_state

    def _axpby(self, x, y, alpha, beta):
        a = (ctypes.c_double * 2)(float(np.real(alpha)), float(np.imag(alpha)))
        b = (ctypes.c_double * 2)(float(np.real(beta)), float(np.imag(beta)))
        call("qj_axpby", _ptr(x), _ptr(y), _NATIVE_CODES[x.dtype], x.numel(), a, b, self._stream())

    def _apply_fanout_gate(self, gate, state, nqubits):
        """FanOut as a CNOT per target (cpu.py:417-431)."""
        control, targets = gate.control_qubits[0], gate.target_qubits
        for target in targets:
            qubits = np.array(sorted([nqubits - control - 1, nqubits - target - 1]), dtype=np.int32)
            state = self._one_qubit_base(state, nqubits, target, GATE_OPS["CNOT"], None, qubits)
        return state

    def _apply_gate(self, gate, state, nqubits):
        """cpu.py:433-450."""
        state = self._as_state(state)
        if gate.name == "fanout":
            return self._apply_fanout_gate(gate, state, nqubits)
        matrix = self._as_custom_matrix(gate)
        qubits = self._create_qubits_tensor(gate, nqubits)
        targets = gate.target_qubits
        if len(targets) == 1:
            op = GATE_OPS.get(gate.__class__.__name__, "apply_gate")
            return self._one_qubit_base(state, nqubits, *targets, op, matrix, qubits)
        if len(targets) == 2:
            op = GATE_OPS.get(gate.__class__.__name__, "apply_two_qubit_gate")
            return self._two_qubit_base(state, nqubits, *targets, op, matrix, qubits)
        return self._multi_qubit_base(state, nqubits, targets, matrix, qubits)

    def _apply_gate_density_matrix(self, gate, state, nqubits, inverse=False):
        """rho -> U rho U^+ as two passes over the 2n-qubit vector (cpu.py:452-498)."""
        name = gate.__class__.__name__
        if name in ["Y", "CY"]:
            return self._apply_ygate_density_matrix(gate, state, nqubits)
        if inverse:
            matrix = np.linalg.inv(np.asarray(self.to_numpy(gate.matrix(self)), dtype=np.complex128))
        else:
            matrix = self._as_custom_matrix(gate)
        qubits = self._create_qubits_tensor(gate, nqubits)
        qubits_dm = qubits + nqubits
        targets = gate.target_qubits
        targets_dm = tuple(q + nqubits for q in targets)
        state = self._as_state(state)
        shape = state.shape
        flat = state.reshape(-1)
        cmatrix = np.conj(matrix) if matrix is not None else None
        if len(targets) == 1:
            op = GATE_OPS.get(name, "apply_gate")
            flat = self._one_qubit_base(flat, 2 * nqubits, *targets, op, matrix, qubits_dm)
            flat = self._one_qubit_base(flat, 2 * nqubits, *targets_dm, op, cmatrix, qubits)
        elif len(targets) == 2:
            op = GATE_OPS.get(name, "apply_two_qubit_gate")
            flat = self._two_qubit_base(flat, 2 * nqubits, *targets, op, matrix, qubits_dm)
            flat = self._two_qubit_base(flat, 2 * nqubits, *targets_dm, op, cmatrix, qubits)
        else:
            flat = self._multi_qubit_base(flat, 2 * nqubits, targets, matrix, qubits_dm)
            flat = self._multi_qubit_base(flat, 2 * nqubits, targets_dm, cmatrix, qubits)
        return flat.reshape(shape)

    def _apply_ygate_density_matrix(self, gate, state, nqubits):
        """Y needs the generic kernel on the column pass to conjugate (cpu.py:500-517)."""
        matrix = self._as_custom_matrix(gate)
        qubits = self._create_qubits_tensor(gate, nqubits)
        qubits_dm = qubits + nqubits
        targets = gate.target_qubits
        targets_dm = tuple(q + nqubits for q in targets)
        state = self._as_state(state)
        shape = state.shape
        flat = state.reshape(-1)
        flat = self._one_qubit_base(flat, 2 * nqubits, *targets, "apply_y", matrix, qubits_dm)
        flat = self._one_qubit_base(flat, 2 * nqubits, *targets_dm, "apply_gate", np.conj(matrix), qubits)
        return flat.reshape(shape)

    def _as_custom_matrix(self, gate, matrices=None):
        """Kernel-facing encoding of ``gate`` (cpu.py:519-539), from
        ``matrices`` (default: the backend's, in its dtype)."""
        if matrices is None:
            matrices = self.custom_matrices
        if isinstance(gate, FusedGate):
            return fused_matrix(lambda g: self._as_custom_matrix(g, matrices), gate)
        params = [self.to_numpy(p) if isinstance(p, torch.Tensor) else p for p in getattr(gate, "parameters", ())]
        return matrices.encode(gate, params)

    @staticmethod
    def _create_qubits_tensor(gate, nqubits):
        """Sorted int32 bit positions of controls and targets (cpu.py:565-569)."""
        qubits = [nqubits - q - 1 for q in gate.control_qubits]
        qubits.extend(nqubits - q - 1 for q in gate.target_qubits)
        return np.array(sorted(qubits), dtype=np.int32)

    def _one_qubit_base(self, state, nqubits, target, kernel, gate, qubits):
        """cpu.py:606-616 -> qj_apply_one_qubit."""
        state = self._as_state(state)
        kind = ONE_QUBIT_KINDS[kernel]
        m = nqubits - target - 1
        mat = None
        if kind in (ONE_QUBIT_KINDS["apply_gate"], ONE_QUBIT_KINDS["apply_z_pow"]):
            mat = self._host_matrix(gate, state.dtype)
        qptr, keep, nq = (None, None, 0) if qubits is None else i32_array(self._seq(qubits))
        call(
            "qj_apply_one_qubit",
            _ptr(state),
            _NATIVE_CODES[state.dtype],
            nqubits,
            kind,
            m,
            qptr,
            nq,
            None if mat is None else mat.ctypes.data_as(ctypes.c_void_p),
            self._stream(),
        )
        del keep
        return state

    def _two_qubit_base(self, state, nqubits, target1, target2, kernel, gate, qubits):
        """cpu.py:618-635 -> qj_apply_two_qubit (first target = matrix MSB)."""
        state = self._as_state(state)
        kind = TWO_QUBIT_KINDS[kernel]
        mat = None
        if kind != TWO_QUBIT_KINDS["apply_swap"]:
            mat = self._host_matrix(gate, state.dtype)
        qptr, keep, nq = (None, None, 0) if qubits is None else i32_array(self._seq(qubits))
        call(
            "qj_apply_two_qubit",
            _ptr(state),
            _NATIVE_CODES[state.dtype],
            nqubits,
            kind,
            nqubits - target1 - 1,
            nqubits - target2 - 1,
            qptr,
            nq,
            None if mat is None else mat.ctypes.data_as(ctypes.c_void_p),
            self._stream(),
        )
        del keep
        return state

    def _multi_qubit_base(self, state, nqubits, targets, gate, qubits):
        """cpu.py:581-604 -> qj_apply_multi_qubit."""
        state = self._as_state(state)
        targets = list(targets)
        if len(targets) > self.MAX_NUM_TARGETS:
            raise ValueError(
                f"Number of target qubits must be <= {self.MAX_NUM_TARGETS} but is {len(targets)}."
            )
        if qubits is None:
            qubits = sorted(nqubits - q - 1 for q in targets)
        mat = self._host_matrix(gate, state.dtype)
        qseq = self._seq(qubits)
        tbits = [nqubits - t - 1 for t in targets]
        if self.mfma and nqubits - len(qseq) >= 4 and (len(targets) == 5 or (len(targets) == 4 and max(tbits) < 8)):
            # dense 5-target block, or 4-target block inside 256-amplitude
            # runs (controlled or not): three real GEMMs on the matrix cores
            # (FP32 MFMA for complex64, FP64 for complex128).  Other 4-target
            # gates stay on the VALU register kernel, which measured as fast
            # or faster there (6.5 vs 7.0 ms at 30 qubits complex128).
            tptr, tkeep, nt = i32_array(tbits)
            qptr, qkeep, nq = i32_array(qseq)
            call("qj_apply_block_mfma", _ptr(state), _NATIVE_CODES[state.dtype], nqubits, tptr, nt,
                 qptr if nq > nt else None, nq, mat.ctypes.data_as(ctypes.c_void_p), self._stream())
            del tkeep, qkeep
            return state
        tbits = [nqubits - t - 1 for t in targets]
        tptr, tkeep, nt = i32_array(tbits)
        qptr, qkeep, nq = i32_array(qseq)
        call(
            "qj_apply_multi_qubit",
            _ptr(state),
            _NATIVE_CODES[state.dtype],
            nqubits,
            tptr,
            nt,
            qptr,
            nq,
            mat.ctypes.data_as(ctypes.c_void_p),
            self._stream(),
        )
        del tkeep, qkeep
        return state

    @staticmethod
    def _seq(qubits):
        if isinstance(qubits, torch.Tensor):
            qubits = qubits.detach().cpu().numpy()
        return [int(q) for q in np.asarray(qubits).reshape(-1)]

    # ------------------------------------------------------------ fusion --
    def matrix_fused(self, fgate):
        """Dense matrix of a FusedGate over its sorted qubits (host, complex128)."""
        return fused_matrix(self._as_custom_matrix, fgate).astype(self.dtype)

    # ------------------------------------------------------- measurement --
    def collapse_state(self, state, qubits, shot, nqubits, normalize=True, density_matrix=False):
        """Project onto outcome ``shot`` of ``qubits`` (gpu.py:616-644, cpu.py:541-563).

        Density matrices (``density_matrix=True`` or a 2-d state) follow qibo's
        density-matrix collapse: entries whose row or column disagrees with
        the outcome are zeroed and, with ``normalize``, the rest is divided by
        its trace.  (The reference's CUDA path ignores ``density_matrix`` and
        runs the state-vector kernel on the flattened matrix.)"""
        if isinstance(shot, torch.Tensor):
            shot = int(shot.reshape(-1)[0])
        elif hasattr(shot, "shape"):
            shot = int(shot.flat[0])
        elif hasattr(shot, "__len__"):
            shot = int(shot[0])
        else:
            shot = int(shot)
        if density_matrix or len(tuple(state.shape)) == 2:
            return self._collapse_density_matrix(state, qubits, shot, nqubits, normalize)
        state = self._as_state(state)
        qptr, keep, nq = i32_array([nqubits - q - 1 for q in reversed(list(qubits))])
        call(
            "qj_collapse_state",
            _ptr(state),
            _NATIVE_CODES[state.dtype],
            nqubits,
            qptr,
            nq,
            shot,
            1 if normalize else 0,
            None,
            self._stream(),
        )
        del keep
        return state

    def _collapse_density_matrix(self, state, qubits, shot, nqubits, normalize=True):
        """rho -> P rho P (/ tr(P rho P)): the 2n-bit collapse keeps the
        measured row bits and column bits equal to ``shot``."""
        state = self._as_state(state)
        qubits = list(qubits)
        k = len(qubits)
        cols = [nqubits - q - 1 for q in reversed(qubits)]
        rows = [2 * nqubits - q - 1 for q in reversed(qubits)]
        qptr, keep, nq = i32_array(cols + rows)
        call("qj_collapse_state", _ptr(state), _NATIVE_CODES[state.dtype], 2 * nqubits, qptr, nq,
             shot | (shot << k), 0, None, self._stream())
        del keep
        if normalize:
            trace = float(self.calculate_probabilities(state, qubits, nqubits, density_matrix=True).sum().item())
            alpha = (ctypes.c_double * 2)(1.0 / trace, 0.0)
            beta = (ctypes.c_double * 2)(0.0, 0.0)
            call("qj_axpby", _ptr(state), _ptr(state), _NATIVE_CODES[state.dtype], state.numel(), alpha, beta,
                 self._stream())
        return state

    def _collapse_statevector(self, state, qubits, shot, nqubits, normalize=True):
        return self.collapse_state(state, qubits, shot, nqubits, normalize)

    def calculate_probabilities(self, state, qubits, nqubits, density_matrix=False):
        """Marginal |psi|^2 over ``qubits`` in list order, FP64 accumulation;
        for a density matrix the marginal of its diagonal."""
        dm = density_matrix or len(tuple(state.shape)) == 2
        state = self._as_state(state)
        rdt = _REAL_OF[state.dtype]
        probs = torch.empty(1 << len(qubits), dtype=rdt, device=self._torch_device)
        qptr, keep, nq = i32_array(list(qubits))
        call(
            "qj_calculate_probabilities_dm" if dm else "qj_calculate_probabilities",
            _ptr(state),
            _NATIVE_CODES[state.dtype],
            nqubits,
            qptr,
            nq,
            _ptr(probs),
            _NATIVE_CODES[rdt],
            self._stream(),
        )
        del keep
        return probs

    def sample_shots(self, probabilities, nshots):
        """Shots drawn on the device (the reference copies the probabilities
        to the host and samples there, gpu.py:770-773): the result of
        ``RandomState.choice(n, nshots, p=probabilities)`` for the backend's
        generator, whose uniforms are drawn here on the host so the stream
        stays the one ``set_seed`` started.  Returns an int64 device tensor."""
        nshots = int(nshots)
        probs = self._as_real_probs(probabilities)
        nstates = probs.numel()
        samples = torch.empty(nshots, dtype=torch.int64, device=self._torch_device)
        if nshots == 0:
            return samples
        rng = getattr(self, "_rng", np.random)  # qibo's backends draw from np.random
        uniforms = torch.from_numpy(rng.random_sample(nshots)).to(self._torch_device)
        call("qj_sample_shots", _ptr(probs), _NATIVE_CODES[probs.dtype], nstates, _ptr(uniforms), nshots,
             _ptr(samples), self._stream())
        return samples

    def calculate_frequencies(self, samples):
        """Counter of sample values (gpu.py:744-748 counts on the host with
        np.unique).  Device int64 samples are counted on the device when the
        value range is small enough for a dense count array."""
        if isinstance(samples, torch.Tensor) and samples.is_cuda and samples.dtype == torch.int64 \
                and samples.numel() > 0:
            nstates = int(samples.max().item()) + 1
            if int(samples.min().item()) >= 0 and nstates <= (1 << 24):
                counts = self._count_samples(samples, nstates)
                return Counter({i: int(c) for i, c in enumerate(counts) if c > 0})
        samples = np.asarray(self.to_numpy(samples))
        res, counts = np.unique(samples, return_counts=True)
        return Counter({int(k): int(v) for k, v in zip(res, counts)})

    def _count_samples(self, samples, nstates):
        freqs = torch.zeros(nstates + 1, dtype=torch.int64, device=self._torch_device)
        bad = freqs[nstates:]
        call("qj_count_samples", _ptr(samples), samples.numel(), _ptr(freqs), nstates, _ptr(bad), self._stream())
        counts = self.to_numpy(freqs)
        if counts[nstates]:
            raise ValueError(f"{int(counts[nstates])} samples outside [0, {nstates}).")
        return counts[:nstates]

    def _as_real_probs(self, probabilities):
        """Probabilities as a contiguous float32/float64 device tensor."""
        probs = probabilities
        if not isinstance(probs, torch.Tensor):
            probs = np.asarray(probs)
            if probs.dtype not in (np.float32, np.float64):
                probs = probs.astype(np.float64)
        probs = self.cast(probs, dtype=_dtype_name(probs.dtype)).reshape(-1).contiguous()
        if probs.dtype not in (torch.float32, torch.float64):
            probs = probs.to(torch.float64)
        return probs

    def sample_frequencies(self, probabilities, nshots):
        """cpu.py:383-394: Metropolis chains from SHOT_METROPOLIS_THRESHOLD shots on;
        below, shots drawn on the device (sample_shots) and counted there."""
        if nshots < SHOT_METROPOLIS_THRESHOLD:
            probs = self._as_real_probs(probabilities)
            samples = self.sample_shots(probs, nshots)
            counts = self._count_samples(samples, probs.numel())
            return Counter({i: int(c) for i, c in enumerate(counts) if c > 0})
        seed = self.random_integers(0, int(1e8), dtype=np.int64)[0]
        nqubits = int(math.log2(tuple(probabilities.shape)[0]))
        frequencies = torch.zeros(2**nqubits, dtype=torch.int64, device=self._torch_device)
        frequencies = self.measure_frequencies_op(
            frequencies, probabilities, nshots, nqubits, seed, self.nthreads
        )
        freqs = self.to_numpy(frequencies)
        return Counter({i: int(f) for i, f in enumerate(freqs) if f > 0})

    def measure_frequencies(self, frequencies, probs, nshots, nqubits, seed, nthreads):
        """Metropolis frequency sampler on the device (ops.py:82-108).

        The chain seeds are drawn on the host exactly as the reference does
        (seed(seed); randint(0, int(1e8)) per chain, NumPy legacy stream); the
        chains themselves, MT19937 included, run in the HIP kernel.  Inputs
        may be host arrays or device tensors; the result has the container
        type and dtype of ``frequencies`` and includes its initial counts.
        """
        del nqubits
        rs = np.random.RandomState(int(seed))
        seeds = np.array([rs.randint(0, int(1e8)) for _ in range(int(nthreads))], dtype=np.uint32)
        host_out = not isinstance(frequencies, torch.Tensor)
        out_dtype = frequencies.dtype
        pdt = _dtype_name(probs.dtype)
        if pdt not in ("float32", "float64"):
            pdt = "float64"
        dprobs = self.cast(probs, dtype=pdt)
        dfreq = self.cast(frequencies, dtype="int64")
        sptr = seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        call(
            "qj_measure_frequencies",
            _ptr(dfreq),
            _ptr(dprobs),
            _NATIVE_CODES[dprobs.dtype],
            dprobs.numel(),
            int(nshots),
            sptr,
            int(nthreads),
            self._stream(),
        )
        if host_out:
            frequencies[...] = self.to_numpy(dfreq).astype(out_dtype)
            return frequencies
        if dfreq.data_ptr() != frequencies.data_ptr():
            frequencies.copy_(dfreq.to(out_dtype))
        return frequencies

    # ----------------------------------------------------- reductions --
    def norm_squared(self, state):
        """sum |psi_i|^2 in FP64."""
        state = self._as_state(state)
        out = ctypes.c_double(0.0)
        call("qj_norm_squared", _ptr(state), _NATIVE_CODES[state.dtype], state.numel(), ctypes.byref(out), self._stream())
        return out.value

    def vdot(self, a, b):
        """<a|b> in FP64."""
        a, b = self._as_state(a), self._as_state(b)
        if a.dtype != b.dtype or a.numel() != b.numel():
            raise ValueError("vdot needs two states of equal size and dtype.")
        out = (ctypes.c_double * 2)()
        call("qj_vdot", _ptr(a), _ptr(b), _NATIVE_CODES[a.dtype], a.numel(), out, self._stream())
        return complex(out[0], out[1])

    def compare_states(self, a, b):
        """FP64 comparison of two states of any complex precision: overlap,
        norms, infidelity 1 - |<a|b>|^2 / (<a|a><b|b>) and max |a_i - b_i|."""
        a, b = self._as_state(a), self._as_state(b)
        if a.numel() != b.numel():
            raise ValueError("compare_states needs two states of equal size.")
        out = (ctypes.c_double * 5)()
        call("qj_compare_states", _ptr(a), _NATIVE_CODES[a.dtype], _ptr(b), _NATIVE_CODES[b.dtype], a.numel(), out,
             self._stream())
        overlap = complex(out[0], out[1])
        return {
            "overlap": overlap,
            "norm2_a": out[2],
            "norm2_b": out[3],
            "infidelity": 1.0 - abs(overlap) ** 2 / (out[2] * out[3]),
            "max_abs_diff": out[4],
        }

    def expectation_tfim(self, state, nqubits, g=1.0):
        """<psi| -sum Z_i Z_{i+1} - g sum X_i |psi> (open chain), FP64."""
        state = self._as_state(state)
        out = ctypes.c_double(0.0)
        call("qj_expectation_tfim", _ptr(state), _NATIVE_CODES[state.dtype], nqubits, float(g), ctypes.byref(out), self._stream())
        return out.value

    # ------------------------------------------------------- execution --
    def execute_circuit(self, circuit, initial_state=None, nshots=1000):
        """Run ``circuit`` from |0...0> or ``initial_state`` (qibo's
        execute_circuit contract, which NumbaBackend inherits).

        With ``self.fusion`` (default), runs of unitary gates go through the
        fused tile executor (a few HBM passes per run instead of one per
        gate); other gates (measurements, channels, wide gates) are applied
        one by one.  A circuit with measurements returns a CircuitResult
        whose frequencies and samples are drawn on the device for ``nshots``
        shots (Metropolis chains from SHOT_METROPOLIS_THRESHOLD shots on);
        one whose measurements collapse the state runs once per shot
        (``execute_circuit_repeated``).  A state the device cannot hold
        raises RuntimeError, as qibo's execute_circuit does.
        """
        if getattr(circuit, "repeated_execution", False):
            return self.execute_circuit_repeated(circuit, nshots, initial_state)
        nqubits = circuit.nqubits
        density_matrix = bool(getattr(circuit, "density_matrix", False))
        try:
            if initial_state is None:
                state = self.zero_state(nqubits, density_matrix=density_matrix)
            else:
                if isinstance(initial_state, QuantumState):
                    initial_state = initial_state.state()
                dtype = _dtype_name(initial_state.dtype) if hasattr(initial_state, "dtype") else self.dtype
                state = self.cast(initial_state, dtype=dtype, copy=True)
            if self.fusion:
                state = self.apply_gates_fused(circuit.queue, state, nqubits)
            else:
                for gate in circuit.queue:
                    if density_matrix:
                        state = gate.apply_density_matrix(self, state, nqubits)
                    else:
                        state = gate.apply(self, state, nqubits)
        except MemoryError:
            raise_error(
                RuntimeError,
                f"State does not fit in {self.device} memory. Please switch the execution device to a "
                "different one using ``qibo.set_device``.",
            )
        if circuit.measurements:
            result = CircuitResult(state, circuit.measurements, backend=self, nshots=nshots)
        else:
            result = QuantumState(state, backend=self)
        circuit._final_state = result  # pylint: disable=protected-access
        return result

    def samples_to_binary(self, samples, nqubits):
        """(nshots, nqubits) int32 0/1 host array, first measured qubit first."""
        samples = np.asarray(self.to_numpy(samples), dtype=np.int64).reshape(-1)
        shifts = np.arange(nqubits - 1, -1, -1, dtype=np.int64)
        return ((samples[:, None] >> shifts) & 1).astype(np.int32)

    def samples_to_decimal(self, samples, nqubits):
        """Outcome integers (int64 host array) of a (nshots, nqubits) 0/1 array."""
        samples = np.asarray(self.to_numpy(samples), dtype=np.int64).reshape(-1, nqubits)
        return samples @ np.left_shift(1, np.arange(nqubits - 1, -1, -1, dtype=np.int64))

    def tile_geometry(self, nbits, tdtype):
        """(low bits, max high bits) of fused tiles: 64 KiB of LDS per tile,
        runs of 2^low contiguous amplitudes (512 B) for coalescing."""
        if tdtype == torch.complex64:
            low, tb = 6, 13
        else:
            low, tb = 5, 12
        low = min(low, nbits)
        return low, max(0, min(tb - low, nbits - low))

    def _custom_matrix_c128(self, gate):
        return self._as_custom_matrix(gate, self._custom_matrices_c128)

    @staticmethod
    def _is_unitary_gate(gate):
        name = gate.__class__.__name__
        return not (name in ("M", "Align", "CallbackGate") or "Channel" in name or hasattr(gate, "coefficients"))

    def apply_gates_fused(self, gates, state, nqubits):
        """Apply ``gates`` in order with the fused tile executor."""
        state = self._as_state(state)
        density_matrix = len(tuple(state.shape)) == 2
        shape = state.shape
        flat = state.reshape(-1)
        nbits = 2 * nqubits if density_matrix else nqubits
        low, max_high = self.tile_geometry(nbits, flat.dtype)
        npdt = np.complex64 if flat.dtype == torch.complex64 else np.complex128
        planner = tiles.Planner(nbits, low, max_high)

        def launch(passes):
            # each pass is launched as soon as the planner closes it, so the
            # host converts the next pass's gates while this one runs
            if passes:
                tiles.run_passes(call, _ptr(flat), _NATIVE_CODES[flat.dtype], nbits, passes, npdt, self._stream())

        for gate in gates:
            seq = tiles.expand_fanout(gate) if gate.name == "fanout" else [gate]
            for g in seq:
                # ops are built from complex128 matrices (exact phase angles)
                # and cast to the state's precision when a pass is encoded
                op = tiles.gate_to_op(self._custom_matrix_c128, g, nbits) if self._is_unitary_gate(g) else None
                if op is None or not tiles.fits_alone(op, low, max_high):
                    launch(planner.finish())
                    st = flat.reshape(shape)
                    st = g.apply_density_matrix(self, st, nqubits) if density_matrix else g.apply(self, st, nqubits)
                    flat = self._as_state(st).reshape(-1)
                    continue
                launch(planner.push(op))
                if density_matrix:
                    launch(planner.push(tiles.column_op(op, nqubits)))
        launch(planner.finish())
        return flat.reshape(shape)

    def assert_allclose(self, value, target, rtol=1e-7, atol=0.0):
        if isinstance(value, QuantumState):
            value = value.state()
        if isinstance(target, QuantumState):
            target = target.state()
        np.testing.assert_allclose(self.to_numpy(value), self.to_numpy(target), rtol=rtol, atol=atol)
