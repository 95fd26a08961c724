"""Parity of the kernels the 30-qubit gate sweep runs, at their strides.

* The sweep's nine gates with its placement (targets 0, n/2, n-1) at 14 and
  22 qubits, against the CPU port of the reference's kernels
  (src/qibojit/custom_operators/gates.py:15-254, 257-335) and the einsum
  closed form.
* Uncontrolled 3-target gates on the three lowest bits (the LDS-staged
  chunk kernel) and one- and two-target gates of every kernel kind with a
  target bit >= 6 (the strided kernels), against the port.
* The full 30-qubit configuration in both precisions, and the 33-qubit
  complex128 / 34-qubit complex64 capacity points, on random product
  states: after a gate on qubits Q (controls, then targets) every amplitude
  is  prod_{q not in Q} a_q[i_q] * (U (x)_{q in Q} a_q)[i_Q],  checked on
  10^4 sampled indices per gate.  Each qubit has its own random factor, so a
  gate applied on a wrong bit changes amplitudes by O(1) of themselves,
  which a U-then-U^dagger round trip cannot detect.
"""

import numpy as np
import pytest

from oracle import einsum_ref as er
from oracle import gate_matrices as gm
from qibojit_amd._qibo import gates
from tests import cases
from tests import product_states as ps
from tests.conftest import ATOL, PARITY

pytestmark = pytest.mark.gpu


def _gate(name, targets, controls, params, rng):
    if name == "Unitary":
        return gates.Unitary(gm.haar_unitary(2 ** len(targets), rng), *targets).controlled_by(*controls)
    if name in ("CNOT", "CU3"):
        return getattr(gates, name)(controls[0], targets[0], *params)
    g = getattr(gates, name)(*targets, *params)
    return g.controlled_by(*controls) if controls else g


def _check(backend, port_backend, gate, n, dtype, rng):
    backend.set_dtype(dtype)
    state = gm.random_state(n, rng, dtype)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state, dtype=dtype, copy=True), n))
    want = port_backend.apply_gate(gate, state.copy(), n)
    np.testing.assert_allclose(got, want, atol=PARITY[dtype], rtol=0)
    ref = er.apply_gate(state, ps.math_matrix(gate), gate.target_qubits, gate.control_qubits, n)
    np.testing.assert_allclose(got, ref, atol=ATOL[dtype])


@pytest.mark.parametrize("n", [14, 22])
@pytest.mark.parametrize("idx", range(27))
def test_sweep_gates_against_port(backend, port_backend, n, idx, dtype, rng):
    label, name, targets, controls, params = cases.sweep_placement(n)[idx]
    gate = _gate(name, targets, controls, params, rng)
    assert gate.target_qubits == tuple(targets), label
    _check(backend, port_backend, gate, n, dtype, rng)


@pytest.mark.parametrize("n,targets", cases.LOW3)
def test_three_lowest_bits(backend, port_backend, n, targets, dtype, rng):
    _check(backend, port_backend, gates.Unitary(gm.haar_unitary(8, rng), *targets), n, dtype, rng)


@pytest.mark.parametrize("n,target,controls", cases.HIGH_BIT_ONE)
@pytest.mark.parametrize("kind", ["Unitary", "X", "Y", "Z", "U1"])
def test_one_target_high_bits(backend, port_backend, n, target, controls, kind, dtype, rng):
    params = (0.917,) if kind == "U1" else ()
    _check(backend, port_backend, _gate(kind, [target], controls, params, rng), n, dtype, rng)


@pytest.mark.parametrize("n,targets,controls", cases.HIGH_BIT_TWO)
@pytest.mark.parametrize("kind", ["Unitary", "SWAP", "fSim", "GeneralizedfSim"])
def test_two_target_high_bits(backend, port_backend, n, targets, controls, kind, dtype, rng):
    params = {"fSim": (0.61, 0.27), "GeneralizedfSim": (gm.haar_unitary(2, rng), 0.35)}.get(kind, ())
    _check(backend, port_backend, _gate(kind, targets, controls, params, rng), n, dtype, rng)


# ------------------------------------------------------ product states --
def _run_product_checks(backend, n, dtype, gate_list, seed, nsamples=10000):
    import torch

    backend.set_dtype(dtype)
    rng = np.random.default_rng(seed)
    factors = ps.factors(n, rng)
    npdt = np.complex64 if dtype == "complex64" else np.complex128
    stored = factors.astype(npdt).astype(np.complex128)
    rtol = 1e-4 if dtype == "complex64" else 1e-10
    for label, gate in gate_list:
        print(f"product-state check n={n} {dtype} {label}", flush=True)
        state = ps.device_product_state(factors.astype(npdt), dtype)
        state = backend.apply_gate(gate, state, n)
        idx = np.concatenate([rng.integers(0, 2**n, size=nsamples), [0, 2**n - 1]]).astype(np.int64)
        got = state[torch.from_numpy(idx).to(state.device)].cpu().numpy().astype(np.complex128)
        want = ps.expected(stored, gate, idx)
        rel = np.abs(got - want) / np.abs(want)
        assert rel.max() <= rtol, (label, float(rel.max()))
        norm = backend.norm_squared(state)
        assert abs(norm - 1) < (1e-4 if dtype == "complex64" else 1e-10), (label, norm)
        del state
        torch.cuda.empty_cache()


def test_sweep_30_qubits_product_state(backend, dtype):
    """The 30-qubit sweep configuration at full size: every sweep gate on a
    fresh random product state, 10^4 amplitudes against the closed form."""
    n = 30
    rng = np.random.default_rng(30)
    gate_list = [(label, _gate(name, t, c, p, rng)) for label, name, t, c, p in cases.sweep_placement(n)]
    _run_product_checks(backend, n, dtype, gate_list, seed=31)


@pytest.mark.parametrize("n,dtype_", [(33, "complex128"), (34, "complex64")])
def test_capacity_product_state(backend, n, dtype_):
    """The largest states one device holds whole (137 GB each)."""
    import torch

    torch.cuda.empty_cache()
    rng = np.random.default_rng(n)
    gate_list = [
        ("H(0)", gates.H(0)),
        (f"U3({n // 2})", gates.U3(n // 2, 0.4, 1.1, -0.6)),
        (f"CNOT({n - 2},{n - 1})", gates.CNOT(n - 2, n - 1)),
        (f"Haar3({n - 3},{n - 2},{n - 1})", gates.Unitary(gm.haar_unitary(8, rng), n - 3, n - 2, n - 1)),
    ]
    _run_product_checks(backend, n, dtype_, gate_list, seed=n + 1, nsamples=4000)
