"""Circuit container, gate fusion grouping and result objects (qibo.models,
qibo.result).

Any fusion grouping is valid: a fused block's dense matrix is the ordered
product of its gates, and blocks keep the order of gates on every qubit.
"""

from collections import Counter

import numpy as np

from qibojit_amd.qibo_compat import gates as gates_module


class Circuit:
    """Ordered list of gates over ``nqubits`` qubits."""

    def __init__(self, nqubits, density_matrix=False):
        self.nqubits = int(nqubits)
        self.density_matrix = density_matrix
        self.queue = []

    def add(self, gate):
        if hasattr(gate, "__iter__") and not isinstance(gate, gates_module.Gate):
            for g in gate:
                self.add(g)
            return
        for q in gate.qubits:
            if q < 0 or q >= self.nqubits:
                raise ValueError(f"Gate on qubit {q} does not fit a {self.nqubits}-qubit circuit.")
        self.queue.append(gate)

    @property
    def measurements(self):
        return [g for g in self.queue if isinstance(g, gates_module.M)]

    @property
    def ngates(self):
        return len(self.queue)

    @property
    def has_collapse(self):
        return any(getattr(g, "collapse", False) for g in self.measurements)

    @property
    def repeated_execution(self):
        """qibo re-runs a circuit once per shot when a measurement collapses
        the state mid-circuit (qibo.models.Circuit.repeated_execution)."""
        return self.has_collapse

    def fuse(self, max_qubits=2):
        """Return a new circuit whose unitary gates are grouped into FusedGates.

        A gate joins the block that is the latest one on each of its qubits
        (qubits no gate has touched yet are free) when the union of qubits
        stays within ``max_qubits``; measurements, channels and FanOut are
        barriers on their qubits.
        """
        barrier = object()
        fused = Circuit(self.nqubits, self.density_matrix)
        last = {}  # qubit -> latest block (or barrier) on it
        for gate in self.queue:
            qs = set(gate.qubits)
            if isinstance(gate, (gates_module.M, gates_module.Channel)) or gate.name == "fanout":
                fused.queue.append(gate)
                for q in qs:
                    last[q] = barrier
                continue
            owners = {last[q] for q in qs if q in last}
            blk = None
            if len(owners) == 1:
                cand = next(iter(owners))
                if cand is not barrier and len(set(cand.target_qubits) | qs) <= max_qubits:
                    blk = cand
            if blk is None:
                blk = gates_module.FusedGate(*sorted(qs))
                fused.queue.append(blk)
            elif not qs <= set(blk.target_qubits):
                newq = tuple(sorted(set(blk.target_qubits) | qs))
                blk.init_args = list(newq)
                blk.target_qubits = newq
            blk.gates.append(gate)
            # only the gate's own qubits move to this block: a later block
            # may already own other qubits of it
            for q in qs:
                last[q] = blk
        # single-gate blocks are emitted as the gate itself
        fused.queue = [
            (g.gates[0] if isinstance(g, gates_module.FusedGate) and len(g.gates) == 1 else g)
            for g in fused.queue
        ]
        return fused


class QuantumState:
    """Final state of an execution, held in the backend's array type."""

    def __init__(self, state, backend):
        self.backend = backend
        self._state = state

    def state(self, numpy=False):
        return self.backend.to_numpy(self._state) if numpy else self._state

    def probabilities(self, qubits=None):
        nqubits = self.backend.nqubits_of(self._state)
        if qubits is None:
            qubits = tuple(range(nqubits))
        return self.backend.calculate_probabilities(self._state, qubits, nqubits)

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.backend.to_numpy(self._state), dtype=dtype)


def frequencies_to_binary(frequencies, nqubits):
    """Counter keyed by bitstrings, the first measured qubit leftmost."""
    return Counter({format(int(k), "b").zfill(nqubits): v for k, v in frequencies.items()})


class MeasurementOutcomes:
    """Samples and frequencies of a circuit's measurements (qibo.result).

    The measured qubits are those of every measurement gate, in gate order;
    outcome k has the first measured qubit as its most significant bit.
    Frequencies come from ``backend.sample_frequencies(probabilities,
    nshots)`` unless samples were drawn first; samples from
    ``backend.sample_shots``."""

    def __init__(self, measurements, backend, probabilities=None, samples=None, nshots=1000):
        self.measurements = list(measurements)
        self.backend = backend
        self.nshots = int(nshots)
        self.measurement_qubits = tuple(q for m in self.measurements for q in m.target_qubits)
        if not self.measurement_qubits:
            raise ValueError("Circuit does not contain measurements.")
        if probabilities is None and samples is None:
            raise ValueError("MeasurementOutcomes needs probabilities or samples.")
        self._probs = probabilities
        self._samples = None if samples is None else np.asarray(samples, dtype=np.int32)
        self._frequencies = None

    def samples(self, binary=True):
        """(nshots, k) 0/1 array, or with ``binary=False`` the outcome integers."""
        k = len(self.measurement_qubits)
        if self._samples is None:
            if self._frequencies is not None:
                # samples that realise the frequencies already reported
                values = np.concatenate([np.repeat(x, f) for x, f in sorted(self._frequencies.items())])
                np.random.shuffle(values)
            else:
                values = self.backend.sample_shots(self._probs, self.nshots)
            self._samples = np.asarray(self.backend.samples_to_binary(values, k), dtype=np.int32)
        if binary:
            return self._samples
        return self.backend.samples_to_decimal(self._samples, k)

    def frequencies(self, binary=True):
        """Counter of outcomes: bitstrings, or integers with ``binary=False``."""
        if self._frequencies is None:
            if self._samples is None:
                self._frequencies = self.backend.sample_frequencies(self._probs, self.nshots)
            else:
                self._frequencies = self.backend.calculate_frequencies(self.samples(binary=False))
        if binary:
            return frequencies_to_binary(self._frequencies, len(self.measurement_qubits))
        return self._frequencies

    def has_samples(self):
        return self._samples is not None


class CircuitResult(QuantumState, MeasurementOutcomes):
    """Final state plus measurement outcomes (qibo.result.CircuitResult):
    the marginal probabilities of the measured qubits are computed once,
    shots are drawn from them on demand."""

    def __init__(self, final_state, measurements, backend=None, samples=None, nshots=1000):
        QuantumState.__init__(self, final_state, backend)
        qubits = tuple(q for m in measurements for q in m.target_qubits)
        if not qubits:
            raise ValueError("Circuit does not contain measurements.")
        probs = QuantumState.probabilities(self, qubits) if samples is None else None
        MeasurementOutcomes.__init__(self, measurements, backend, probabilities=probs, samples=samples,
                                     nshots=nshots)

    def probabilities(self, qubits=None):
        if qubits is None:
            qubits = self.measurement_qubits
        return QuantumState.probabilities(self, qubits)
