"""Matrix-core 4- and 5-qubit blocks (FP32 MFMA on complex64, FP64 MFMA on
complex128, with and without controls) and the block fusion, against the
VALU kernel, the CPU port and closed forms."""

import numpy as np
import pytest

from oracle import einsum_ref
from oracle.gate_matrices import haar_unitary, random_state

pytestmark = pytest.mark.gpu


TOL = {"complex64": 1e-5, "complex128": 1e-12}


@pytest.mark.parametrize("n", [10, 12, 15])
@pytest.mark.parametrize("targets", [(0, 1, 2, 3, 4), (4, 2, 0, 3, 1), (3, 7, 9, 1, 5), (9, 8, 7, 6, 5),
                                     (8, 1, 5, 0, 3), (6, 4, 2, 7, 5)])
def test_mfma_block_matches_oracle(backend, rng, n, targets, dtype):
    """5-target blocks on matrix cores (FP32 MFMA for complex64, FP64 for
    complex128) against the einsum oracle and the VALU register kernel."""
    from qibojit_amd._qibo import gates

    if max(targets) >= n:
        pytest.skip("targets outside the state")
    backend.set_dtype(dtype)
    u = haar_unitary(32, rng)
    state = random_state(n, rng).astype(dtype)
    gate = gates.Unitary(u, *targets)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state.copy()), n))
    want = einsum_ref.apply_gate(state.astype(np.complex128), u, list(targets), nqubits=n)
    assert np.max(np.abs(got - want)) < TOL[dtype]
    # the VALU register kernel agrees
    backend.mfma = False
    try:
        valu = backend.to_numpy(backend.apply_gate(gate, backend.cast(state.copy()), n))
    finally:
        backend.mfma = True
    assert np.max(np.abs(got - valu)) < TOL[dtype]


@pytest.mark.parametrize("n,targets,controls", [
    (12, (7, 8, 9, 10, 11), (0,)),          # staged block, control on the top bit
    (12, (11, 10, 8, 9), (1, 2)),           # 4 targets, two high controls
    (13, (0, 1, 2, 3, 4), (12,)),           # control on bit 0, targets on top
    (12, (3, 9, 5, 11), (10,)),             # control between targets, below bit 8
    (14, (13, 12, 11, 10, 9), (5, 0, 2)),   # three controls
    (13, (6, 7, 8, 9), (12, 11)),           # controls inside the staged bits
])
def test_controlled_blocks_on_matrix_cores(backend, port_backend, rng, n, targets, controls, dtype):
    """Controlled 4/5-target blocks on matrix cores against the CPU port of
    the reference's multicontrol kernels (gates.py:265-411)."""
    from qibojit_amd._qibo import gates

    backend.set_dtype(dtype)
    u = haar_unitary(2 ** len(targets), rng)
    state = random_state(n, rng).astype(dtype)
    gate = gates.Unitary(u, *targets).controlled_by(*controls)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state.copy()), n))
    want = port_backend.apply_gate(gate, state.astype(np.complex128), n)
    assert np.max(np.abs(got - want)) < TOL[dtype]


def test_mfma_entry_rejects_bad_arguments(backend):
    import ctypes

    from qibojit_amd import _native

    # real buffers sized for every call below (2^12 amplitudes, a 32x32
    # matrix), so a call that slipped past validation would stay in bounds
    backend.set_dtype("complex128")
    state = backend.zero_state(12)
    sp = ctypes.c_void_p(state.data_ptr())
    mat = np.eye(32, dtype=np.complex128)
    mp = mat.ctypes.data_as(ctypes.c_void_p)
    s = backend._stream()
    c128 = _native.QJ_COMPLEX128

    def call(dtype, n, targets, qubits):
        t, tk, nt = _native.i32_array(targets)
        if qubits is None:
            _native.call("qj_apply_block_mfma", sp, dtype, n, t, nt, None, 0, mp, s)
        else:
            q, qk, nq = _native.i32_array(qubits)
            _native.call("qj_apply_block_mfma", sp, dtype, n, t, nt, q, nq, mp, s)

    with pytest.raises(ValueError, match="dtype"):
        call(_native.QJ_FLOAT64, 12, [0, 1, 2, 3, 4], None)
    with pytest.raises(ValueError, match="4 free qubits"):
        call(c128, 8, [0, 1, 2, 3, 4], None)
    with pytest.raises(ValueError, match="4 or 5 targets"):
        call(c128, 12, [0, 1, 2], None)
    with pytest.raises(ValueError, match="repeated"):
        call(c128, 12, [0, 1, 2, 3, 3], None)
    with pytest.raises(ValueError, match="out of range"):
        call(c128, 12, [0, 1, 2, 3, 12], None)
    with pytest.raises(ValueError, match="ascending"):
        call(c128, 12, [0, 1, 2, 3], [0, 1, 3, 2, 9])
    with pytest.raises(ValueError, match="every target"):
        call(c128, 12, [0, 1, 2, 3], [0, 1, 2, 8, 9])
    with pytest.raises(ValueError, match="4 free qubits"):
        call(c128, 12, [0, 1, 2, 3, 4], [0, 1, 2, 3, 4, 5, 6, 7, 8])
    # nothing was launched: the state is still |0...0>
    assert backend.to_numpy(state)[0] == 1


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


def _block_direct(backend, state, n, targets, controls, u):
    """qj_apply_block_mfma called through the C-ABI (whatever the dispatch
    of HipBackend would pick for this placement)."""
    import ctypes

    from qibojit_amd import _native
    from qibojit_amd.backends.hip import _NATIVE_CODES

    tbits = [n - 1 - t for t in targets]
    qubits = sorted(tbits + [n - 1 - c for c in controls])
    t, tk, nt = _native.i32_array(tbits)
    q, qk, nq = _native.i32_array(qubits)
    mat = np.ascontiguousarray(u.astype(np.complex64 if state.dtype == np.complex64 else np.complex128))
    dev = backend.cast(state.copy())
    _native.call("qj_apply_block_mfma", ctypes.c_void_p(dev.data_ptr()), _NATIVE_CODES[dev.dtype], n, t, nt,
                 q if controls else None, nq, mat.ctypes.data_as(ctypes.c_void_p), backend._stream())
    return backend.to_numpy(dev)


@pytest.mark.parametrize("targets", [(0, 1, 2, 3), (5, 2, 9, 7), (8, 9, 10, 11), (1, 2, 3, 4), (11, 6, 10, 7)])
@pytest.mark.parametrize("controls", [(), (6,)])
def test_four_qubit_blocks_on_matrix_cores(backend, port_backend, rng, targets, controls, dtype):
    """4-qubit blocks as 16x16 products on the matrix cores, called directly
    for every placement, and through the backend's dispatch (matrix cores
    when the targets lie below bit 8, the VALU kernel otherwise), against the
    CPU port of the reference's kernels."""
    from qibojit_amd._qibo import gates

    if set(controls) & set(targets):
        pytest.skip("control among the targets")
    n = 12
    backend.set_dtype(dtype)
    u = haar_unitary(16, rng)
    state = random_state(n, rng).astype(dtype)
    gate = gates.Unitary(u, *targets).controlled_by(*controls)
    want = port_backend.apply_gate(gate, state.astype(np.complex128), n)
    got = backend.to_numpy(backend.apply_gate(gate, backend.cast(state.copy()), n))
    assert np.max(np.abs(got - want)) < TOL[dtype]
    direct = _block_direct(backend, state, n, targets, controls, u)
    assert np.max(np.abs(direct - want)) < TOL[dtype]
