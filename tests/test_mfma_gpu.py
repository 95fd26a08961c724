"""FP32 matrix-core 5-qubit blocks and the block fusion, against the VALU
kernel, the CPU port and closed forms."""

import numpy as np
import pytest

from oracle import einsum_ref
from oracle.gate_matrices import haar_unitary, random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [10, 12, 15])
@pytest.mark.parametrize("targets", [(0, 1, 2, 3, 4), (4, 2, 0, 3, 1), (3, 7, 9, 1, 5), (9, 8, 7, 6, 5),
                                     (8, 1, 5, 0, 3), (6, 4, 2, 7, 5)])
def test_mfma_block_matches_oracle(backend, rng, n, targets):
    from qibojit_amd._qibo import gates

    if max(targets) >= n:
        pytest.skip("targets outside the state")
    backend.set_dtype("complex64")
    u = haar_unitary(32, rng)
    state = random_state(n, rng).astype(np.complex64)
    gate = gates.Unitary(u, *targets)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state.copy()), n))
    want = einsum_ref.apply_gate(state.astype(np.complex128), u, list(targets), nqubits=n)
    assert np.max(np.abs(got - want)) < 1e-5
    # the VALU register kernel agrees
    backend.mfma = False
    try:
        valu = backend.to_numpy(backend.apply_gate(gate, backend.cast(state.copy()), n))
    finally:
        backend.mfma = True
    assert np.max(np.abs(got - valu)) < 1e-5


def test_mfma_entry_rejects_bad_arguments(backend):
    import ctypes

    from qibojit_amd import _native

    # real buffers sized for every call below (2^12 amplitudes, a 32x32
    # matrix), so a call that slipped past validation would stay in bounds
    backend.set_dtype("complex64")
    state = backend.zero_state(12)
    sp = ctypes.c_void_p(state.data_ptr())
    mat = np.eye(32, dtype=np.complex64)
    mp = mat.ctypes.data_as(ctypes.c_void_p)
    s = backend._stream()
    t, keep, nt = _native.i32_array([0, 1, 2, 3, 4])
    with pytest.raises(ValueError, match="complex64"):
        _native.call("qj_apply_block_mfma", sp, _native.QJ_COMPLEX128, 11, t, nt, mp, s)
    with pytest.raises(ValueError, match="10..40"):
        _native.call("qj_apply_block_mfma", sp, _native.QJ_COMPLEX64, 9, t, nt, mp, s)
    t4, keep4, n4 = _native.i32_array([0, 1, 2, 3])
    with pytest.raises(ValueError, match="exactly 5"):
        _native.call("qj_apply_block_mfma", sp, _native.QJ_COMPLEX64, 12, t4, n4, mp, s)
    tr, keepr, nr = _native.i32_array([0, 1, 2, 3, 3])
    with pytest.raises(ValueError, match="repeated"):
        _native.call("qj_apply_block_mfma", sp, _native.QJ_COMPLEX64, 12, tr, nr, mp, s)
    backend.set_dtype("complex128")


@pytest.mark.parametrize("dtype_", ["complex64", "complex128"])
def test_fused_brickwork_matches_unfused(backend, port_backend, rng, dtype_):
    from qibojit_amd._qibo import Circuit, gates
    from qibojit_amd.fusion import fuse_blocks

    n, depth = 12, 8
    c = Circuit(n)
    for d in range(depth):
        for q in range(d % 2, n - 1, 2):
            c.add(gates.Unitary(haar_unitary(4, rng), q, q + 1))
    fused = fuse_blocks(c, 5)
    assert len(fused.queue) < len(c.queue)
    backend.set_dtype(dtype_)
    state = backend.zero_state(n)
    for g in fused.queue:
        state = backend.apply_gate(g, state, n)
    got = backend.to_numpy(state)
    want = np.zeros(2**n, dtype=np.complex128)
    want[0] = 1
    for g in c.queue:
        want = port_backend.apply_gate(g, want, n)
    tol = 1e-5 if dtype_ == "complex64" else 1e-12
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("targets", [(0, 1, 2, 3), (5, 2, 9, 7), (8, 9, 10, 11), (1, 2, 3, 4), (11, 6, 10, 7)])
def test_four_qubit_blocks_on_matrix_cores(backend, rng, targets):
    """A 4-qubit block runs as U x I with a spectator qubit on the MFMA kernel."""
    from qibojit_amd._qibo import gates

    n = 12
    backend.set_dtype("complex64")
    u = haar_unitary(16, rng)
    state = random_state(n, rng).astype(np.complex64)
    gate = gates.Unitary(u, *targets)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state.copy()), n))
    want = einsum_ref.apply_gate(state.astype(np.complex128), u, list(targets), nqubits=n)
    assert np.max(np.abs(got - want)) < 1e-5
