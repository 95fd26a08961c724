"""Single import point for the qibo interface.

Uses qibo itself when it is importable (the drop-in case: qibo 0.3.x loads
this backend through its plugin registry), otherwise the stand-ins of
:mod:`qibojit_amd.qibo_compat`, which expose the same names.
"""

try:  # pragma: no cover - exercised only where qibo is installed
    from qibo import Circuit, gates
    # qibo 0.3's Backend carries execute_circuit_repeated and the sample
    # conversions; qibojit's backends derive from it (cpu.py:9, :39)
    from qibo.backends import Backend, NumpyMatrices
    from qibo.config import SHOT_METROPOLIS_THRESHOLD, raise_error
    from qibo.gates.abstract import Gate, ParametrizedGate
    from qibo.gates.special import FusedGate
    from qibo.result import CircuitResult, MeasurementOutcomes, QuantumState

    HAVE_QIBO = True
except ImportError:
    from qibojit_amd.qibo_compat import gates
    from qibojit_amd.qibo_compat.backend import Backend
    from qibojit_amd.qibo_compat.config import SHOT_METROPOLIS_THRESHOLD, raise_error
    from qibojit_amd.qibo_compat.gates import FusedGate, Gate, ParametrizedGate
    from qibojit_amd.qibo_compat.models import Circuit, CircuitResult, MeasurementOutcomes, QuantumState
    from qibojit_amd.qibo_compat.npmatrices import NumpyMatrices

    HAVE_QIBO = False

__all__ = [
    "Backend",
    "Circuit",
    "CircuitResult",
    "FusedGate",
    "Gate",
    "HAVE_QIBO",
    "MeasurementOutcomes",
    "NumpyMatrices",
    "ParametrizedGate",
    "QuantumState",
    "SHOT_METROPOLIS_THRESHOLD",
    "gates",
    "raise_error",
]
