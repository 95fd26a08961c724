"""What the gate kernels read for each gate.

The kernels do not always take a gate's full matrix (the encodings of
src/qibojit/backends/matrices.py:7-70, which the dispatch contract fixes):

* a controlled gate arrives as the block its controls select, since the
  controls only choose which amplitudes the kernel visits;
* U1 / CU1 arrive as the scalar e^{i theta} the z_pow kernel multiplies the
  |1> half by;
* fSim / GeneralizedfSim arrive as five numbers, the {|01>, |10>} block row
  by row and then the |11> phase e^{-i phi}, which the fsim kernel reads;
* DEUTSCH arrives as i RX(2 theta), its block under two controls.

Every other gate arrives as its textbook matrix over its targets.  The
encodings are host arrays: the library copies them into kernel arguments
or its staging ring.
"""

import numpy as np

from qibojit_amd._qibo import NumpyMatrices

# controlled gate -> the uncontrolled gate whose matrix is its selected block
SELECTED_BLOCK = {
    "CNOT": "X",
    "TOFFOLI": "X",
    "CY": "Y",
    "CZ": "Z",
    "CCZ": "Z",
    "CH": "H",
    "CSX": "SX",
    "CSXDG": "SXDG",
    "CRX": "RX",
    "CRY": "RY",
    "CRZ": "RZ",
    "CU1": "U1",
    "CU2": "U2",
    "CU3": "U3",
}


def _z_pow_scalar(theta, dtype):
    return np.asarray(np.exp(1j * theta), dtype=dtype)


def _fsim_five(block, phi, dtype):
    entries = list(np.asarray(block, dtype=np.complex128).reshape(4)) + [np.exp(-1j * phi)]
    return np.asarray(entries, dtype=dtype)


def _fsim_block(theta):
    c, s = np.cos(theta), -1j * np.sin(theta)
    return [[c, s], [s, c]]


class KernelMatrices:
    """Kernel-facing encodings of gates in one precision ("complex64" or
    "complex128"), built on the textbook matrices of ``NumpyMatrices``."""

    def __init__(self, dtype):
        self.dtype = str(np.dtype(dtype))
        self.textbook = NumpyMatrices(self.dtype)

    def encode(self, gate, parameters=None):
        """The array the kernels read for ``gate``; ``parameters`` (host
        numbers) default to ``gate.parameters``."""
        name = gate.__class__.__name__
        params = list(getattr(gate, "parameters", ()) if parameters is None else parameters)
        kind = SELECTED_BLOCK.get(name, name)
        if kind == "U1":
            return _z_pow_scalar(params[0], self.dtype)
        if kind == "fSim":
            return _fsim_five(_fsim_block(params[0]), params[1], self.dtype)
        if kind == "GeneralizedfSim":
            return _fsim_five(params[0], params[1], self.dtype)
        if kind == "DEUTSCH":
            return np.asarray(1j * self.textbook.RX(2 * params[0]), dtype=self.dtype)
        form = getattr(self.textbook, kind)
        if kind == "FanOut":
            return form(*gate.init_args)
        if kind == "GeneralizedRBS":
            return form(gate.init_args[0], gate.init_args[1], gate.init_kwargs["theta"], gate.init_kwargs["phi"])
        if not callable(form):
            return form
        if params:
            return form(*params)
        # parameter-free families sized by their targets (identity, ...)
        return form(2 ** len(gate.target_qubits))
