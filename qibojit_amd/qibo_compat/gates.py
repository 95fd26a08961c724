"""Gate descriptors with the fields qibojit's backends read from qibo gates.

A backend consumes ``gate.__class__.__name__``, ``gate.name``,
``gate.target_qubits``, ``gate.control_qubits``, ``gate.parameters``,
``gate.init_args``, ``gate.init_kwargs`` and ``gate.matrix(backend)``
(src/qibojit/backends/cpu.py:417-539).  The classes below carry exactly
those, with qibo's argument order for each constructor.
"""

import numpy as np

from qibojit_amd.qibo_compat.npmatrices import NumpyMatrices


class Gate:
    """Base gate: a name, target and control qubits."""

    def __init__(self):
        self.name = None
        self.init_args = []
        self.init_kwargs = {}
        self.target_qubits = ()
        self.control_qubits = ()
        self.is_controlled_by = False
        self._parameters = ()

    @property
    def parameters(self):
        return self._parameters

    @property
    def qubits(self):
        return tuple(self.control_qubits) + tuple(self.target_qubits)

    def _set_targets_and_controls(self, targets, controls=()):
        targets, controls = tuple(targets), tuple(controls)
        all_q = targets + controls
        if len(set(all_q)) != len(all_q):
            raise ValueError(f"Gate {self.name} acts twice on one qubit: {all_q}.")
        self.target_qubits = targets
        self.control_qubits = controls

    def controlled_by(self, *qubits):
        """Add control qubits (qibo: Gate.controlled_by)."""
        if qubits:
            self.is_controlled_by = True
            self._set_targets_and_controls(self.target_qubits, tuple(self.control_qubits) + tuple(qubits))
        return self

    def matrix(self, backend=None):
        """Dense matrix of the gate in qibo's convention (over its qubits)."""
        dtype = getattr(backend, "dtype", "complex128") if backend is not None else "complex128"
        matrices = NumpyMatrices(dtype)
        _matrix = getattr(matrices, self.__class__.__name__)
        if self.__class__.__name__ == "FanOut":
            return _matrix(*self.init_args)
        if isinstance(self, ParametrizedGate):
            return _matrix(*self.parameters)
        return _matrix(2 ** len(self.target_qubits)) if callable(_matrix) else _matrix

    def apply(self, backend, state, nqubits):
        return backend.apply_gate(self, state, nqubits)

    def apply_density_matrix(self, backend, state, nqubits):
        return backend.apply_gate_density_matrix(self, state, nqubits)


class ParametrizedGate(Gate):
    """Gate whose matrix depends on ``parameters``."""

    def __init__(self, parameters):
        super().__init__()
        self._parameters = tuple(parameters)

    @property
    def parameters(self):
        return self._parameters

    @parameters.setter
    def parameters(self, values):
        if not isinstance(values, (tuple, list)):
            values = (values,)
        self._parameters = tuple(values)


class _Fixed1(Gate):
    def __init__(self, q):
        super().__init__()
        self.name = self.__class__.__name__.lower()
        self.init_args = [q]
        self._set_targets_and_controls((q,))


class H(_Fixed1):
    pass


class X(_Fixed1):
    pass


class Y(_Fixed1):
    pass


class Z(_Fixed1):
    pass


class S(_Fixed1):
    pass


class SDG(_Fixed1):
    pass


class T(_Fixed1):
    pass


class TDG(_Fixed1):
    pass


class SX(_Fixed1):
    pass


class SXDG(_Fixed1):
    pass


class I(Gate):  # noqa: E742
    def __init__(self, *q):
        super().__init__()
        self.name = "id"
        self.init_args = list(q)
        self._set_targets_and_controls(q)


class _Param1(ParametrizedGate):
    nparams = 1

    def __init__(self, q, *params, **kwargs):
        names = self._param_names
        values = list(params) + [kwargs[k] for k in names[len(params):]]
        super().__init__(values)
        self.name = self.__class__.__name__.lower()
        self.init_args = [q]
        self.init_kwargs = dict(zip(names, values))
        self._set_targets_and_controls((q,))


class RX(_Param1):
    _param_names = ("theta",)


class RY(_Param1):
    _param_names = ("theta",)


class RZ(_Param1):
    _param_names = ("theta",)


class U1(_Param1):
    _param_names = ("theta",)


class U2(_Param1):
    _param_names = ("phi", "lam")


class U3(_Param1):
    _param_names = ("theta", "phi", "lam")


class _Fixed2c(Gate):
    """Named controlled gate: first qubit control, second target."""

    def __init__(self, q0, q1):
        super().__init__()
        self.name = self.__class__.__name__.lower()
        self.init_args = [q0, q1]
        self._set_targets_and_controls((q1,), (q0,))


class CNOT(_Fixed2c):
    pass


class CY(_Fixed2c):
    pass


class CZ(_Fixed2c):
    pass


class CSX(_Fixed2c):
    pass


class CSXDG(_Fixed2c):
    pass


class CH(_Fixed2c):
    pass


class _Param2c(ParametrizedGate):
    def __init__(self, q0, q1, *params, **kwargs):
        names = self._param_names
        values = list(params) + [kwargs[k] for k in names[len(params):]]
        super().__init__(values)
        self.name = self.__class__.__name__.lower()
        self.init_args = [q0, q1]
        self.init_kwargs = dict(zip(names, values))
        self._set_targets_and_controls((q1,), (q0,))


class CRX(_Param2c):
    _param_names = ("theta",)


class CRY(_Param2c):
    _param_names = ("theta",)


class CRZ(_Param2c):
    _param_names = ("theta",)


class CU1(_Param2c):
    _param_names = ("theta",)


class CU2(_Param2c):
    _param_names = ("phi", "lam")


class CU3(_Param2c):
    _param_names = ("theta", "phi", "lam")


class _Fixed2(Gate):
    def __init__(self, q0, q1):
        super().__init__()
        self.name = self.__class__.__name__.lower()
        self.init_args = [q0, q1]
        self._set_targets_and_controls((q0, q1))


class SWAP(_Fixed2):
    pass


class iSWAP(_Fixed2):
    pass


class FSWAP(_Fixed2):
    pass


class _Param2(ParametrizedGate):
    def __init__(self, q0, q1, *params, **kwargs):
        names = self._param_names
        values = list(params) + [kwargs[k] for k in names[len(params):]]
        super().__init__(values)
        self.name = self.__class__.__name__.lower()
        self.init_args = [q0, q1]
        self.init_kwargs = dict(zip(names, values))
        self._set_targets_and_controls((q0, q1))


class fSim(_Param2):
    _param_names = ("theta", "phi")


class GeneralizedfSim(_Param2):
    _param_names = ("unitary", "phi")

    def __init__(self, q0, q1, unitary, phi):
        super().__init__(q0, q1, np.asarray(unitary), phi)
        self.name = "generalizedfsim"


class RXX(_Param2):
    _param_names = ("theta",)


class RYY(_Param2):
    _param_names = ("theta",)


class RZZ(_Param2):
    _param_names = ("theta",)


class TOFFOLI(Gate):
    def __init__(self, q0, q1, q2):
        super().__init__()
        self.name = "ccx"
        self.init_args = [q0, q1, q2]
        self._set_targets_and_controls((q2,), (q0, q1))


class CCZ(Gate):
    def __init__(self, q0, q1, q2):
        super().__init__()
        self.name = "ccz"
        self.init_args = [q0, q1, q2]
        self._set_targets_and_controls((q2,), (q0, q1))


class DEUTSCH(ParametrizedGate):
    def __init__(self, q0, q1, q2, theta):
        super().__init__([theta])
        self.name = "deutsch"
        self.init_args = [q0, q1, q2]
        self.init_kwargs = {"theta": theta}
        self._set_targets_and_controls((q2,), (q0, q1))


class FanOut(Gate):
    """CNOT from q[0] onto each of q[1:]."""

    def __init__(self, *q):
        super().__init__()
        self.name = "fanout"
        self.init_args = list(q)
        self._set_targets_and_controls(q[1:], (q[0],))


class Unitary(ParametrizedGate):
    def __init__(self, unitary, *q, trainable=True, name=None, check_unitary=False):
        super().__init__([unitary])
        self.name = name or "Unitary"
        self.init_args = [unitary] + list(q)
        self.init_kwargs = {"trainable": trainable}
        self._set_targets_and_controls(q)
        m = np.asarray(unitary)
        if m.shape != (2 ** len(q), 2 ** len(q)):
            raise ValueError(f"Unitary of shape {m.shape} does not act on {len(q)} qubits.")


class FusedGate(Gate):
    """Several gates merged into one dense block over their qubits."""

    def __init__(self, *q):
        super().__init__()
        self.name = "fused"
        self.init_args = list(q)
        self.gates = []
        self._set_targets_and_controls(tuple(sorted(q)))

    @classmethod
    def from_gate(cls, gate):
        fgate = cls(*gate.qubits)
        fgate.append(gate)
        return fgate

    def append(self, gate):
        if not set(gate.qubits) <= set(self.target_qubits):
            raise ValueError("Gate does not act on the qubits of the fused block.")
        self.gates.append(gate)


class M(Gate):
    """Measurement of ``q``.  With ``collapse=True`` the state is projected
    on an outcome drawn from the backend's generator (qibo.gates.M.apply:
    probabilities of the sorted qubits, one shot, collapse); otherwise the
    gate leaves the state as it is and the circuit's result samples it."""

    def __init__(self, *q, collapse=False):
        super().__init__()
        self.name = "measure"
        self.init_args = list(q)
        self.collapse = collapse
        self.drawn = []  # outcomes of the collapsing applications, in order
        self._set_targets_and_controls(q)

    def apply(self, backend, state, nqubits):
        if not self.collapse:
            return state
        qubits = sorted(self.target_qubits)
        probs = backend.calculate_probabilities(state, qubits, nqubits)
        shot = int(np.asarray(backend.to_numpy(backend.sample_shots(probs, 1))).reshape(-1)[0])
        self.drawn.append(shot)
        return backend.collapse_state(state, qubits, shot, nqubits)

    def apply_density_matrix(self, backend, state, nqubits):
        return self.apply(backend, state, nqubits)


class Channel(Gate):
    """Probabilistic mixture of gates (qibo.gates.channels)."""

    def apply(self, backend, state, nqubits):
        return backend.apply_channel(self, state, nqubits)

    def apply_density_matrix(self, backend, state, nqubits):
        return backend.apply_channel(self, state, nqubits)


class KrausChannel(Channel):
    """rho -> sum_i K_i rho K_i^dagger for Kraus operators ``(qubits_i, K_i)``
    (qibo.gates.channels.KrausChannel): coefficients 1, not unitary, so the
    backend applies each K_i to a copy of rho."""

    def __init__(self, qubits, operators):
        super().__init__()
        self.name = "KrausChannel"
        if isinstance(qubits, int) or (len(qubits) and isinstance(qubits[0], (int, np.integer))):
            qubits = [tuple(qubits) if not isinstance(qubits, int) else (qubits,)] * len(operators)
        self.init_args = [qubits, operators]
        self.gates = tuple(Unitary(np.asarray(m), *q) for q, m in zip(qubits, operators))
        self.coefficients = len(self.gates) * (1.0,)
        self.coefficient_sum = 1.0
        self._all_unitary_operators = False
        all_q = sorted({q for qs in qubits for q in qs})
        self.target_qubits = tuple(all_q)


class ReadoutErrorChannel(KrausChannel):
    """Readout error with row-stochastic matrix P on ``qubits``
    (qibo.gates.channels.ReadoutErrorChannel): Kraus operators
    K_jk = sqrt(P[j, k]) |k><j|."""

    def __init__(self, qubits, probabilities):
        qubits = (qubits,) if isinstance(qubits, int) else tuple(qubits)
        p = np.asarray(probabilities, dtype=float)
        d = 2 ** len(qubits)
        if p.shape != (d, d):
            raise ValueError(f"Readout matrix of shape {p.shape} does not act on {len(qubits)} qubits.")
        if np.any(p < 0) or not np.allclose(p.sum(axis=1), 1):
            raise ValueError("The readout matrix must be row-stochastic.")
        ops = []
        for j in range(d):
            for k in range(d):
                op = np.zeros((d, d), dtype=complex)
                op[k, j] = np.sqrt(p[j, k])
                ops.append(op)
        super().__init__(qubits, ops)
        self.name = "ReadoutErrorChannel"
        self.init_args = [qubits, probabilities]


class UnitaryChannel(Channel):
    """sum_i p_i U_i rho U_i^dagger + (1 - sum p_i) rho."""

    def __init__(self, qubits, operators):
        super().__init__()
        self.name = "UnitaryChannel"
        self.init_args = [qubits, operators]
        self.coefficients = tuple(p for p, _ in operators)
        self.gates = tuple(Unitary(m, *q) for q, (_, m) in zip(qubits, operators))
        self.coefficient_sum = float(sum(self.coefficients))
        self._all_unitary_operators = True
        all_q = sorted({q for qs in qubits for q in qs})
        self.target_qubits = tuple(all_q)
