"""Host-side logic and the C-ABI boundary, on CPU (no GPU compute calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

import qibojit_amd
from qibojit_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qjhip.h")


def _declared_symbols():
    text = open(HEADER, encoding="utf-8").read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int64_t|int)\s+(qj_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    syms = _declared_symbols()
    assert "qj_apply_one_qubit" in syms and "qj_measure_frequencies" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol():
    lib = _native.lib()  # loads without a GPU
    for name in _declared_symbols():
        assert hasattr(lib, name), name
    exported = set(_native.SIGNATURES) | set(_native.VALUE_SIGNATURES) | {"qj_last_error"}
    assert exported == set(_declared_symbols())


def test_abi_version_and_device_count():
    assert _native.lib().qj_abi_version() == _native.ABI_VERSION
    if not os.path.exists("/dev/kfd"):
        assert _native.device_count() == 0


def _fake_ptr():
    return ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first


def test_argument_errors_raise_before_any_launch():
    qptr, keep, nq = _native.i32_array([3, 1])  # unsorted qubits tensor
    with pytest.raises(ValueError, match="sorted"):
        _native.call("qj_apply_one_qubit", _fake_ptr(), _native.QJ_COMPLEX128, 5, 0, 1, qptr, nq, None, None)
    with pytest.raises(ValueError, match="out of range"):
        _native.call("qj_apply_one_qubit", _fake_ptr(), _native.QJ_COMPLEX128, 5, 1, 7, None, 0, None, None)
    with pytest.raises(ValueError, match="missing"):
        q2, k2, n2 = _native.i32_array([0, 2])
        _native.call("qj_apply_one_qubit", _fake_ptr(), _native.QJ_COMPLEX128, 5, 1, 1, q2, n2, None, None)
    with pytest.raises(ValueError, match="<= 12"):
        t, kt, nt = _native.i32_array(list(range(13)))
        _native.call("qj_apply_multi_qubit", _fake_ptr(), _native.QJ_COMPLEX64, 14, t, nt, None, 0, _fake_ptr(), None)
    with pytest.raises(ValueError, match="dtype"):
        _native.call("qj_apply_one_qubit", _fake_ptr(), _native.QJ_FLOAT32, 5, 1, 1, None, 0, None, None)
    with pytest.raises(ValueError, match="null state"):
        _native.call("qj_apply_two_qubit", None, _native.QJ_COMPLEX64, 5, 1, 1, 2, None, 0, None, None)
    with pytest.raises(ValueError, match="unique|repeated"):
        _native.call("qj_apply_two_qubit", _fake_ptr(), _native.QJ_COMPLEX64, 5, 1, 2, 2, None, 0, None, None)
    with pytest.raises(ValueError, match="result"):
        qk, kk, nk = _native.i32_array([0, 1])
        _native.call("qj_collapse_state", _fake_ptr(), _native.QJ_COMPLEX64, 4, qk, nk, 4, 1, None, None)
    del keep


def test_gate_ops_table_matches_reference():
    from qibojit_amd.backends.hip import GATE_OPS

    assert GATE_OPS == {
        "X": "apply_x", "CNOT": "apply_x", "TOFFOLI": "apply_x", "Y": "apply_y", "Z": "apply_z",
        "CY": "apply_y", "CZ": "apply_z", "U1": "apply_z_pow", "CU1": "apply_z_pow",
        "SWAP": "apply_swap", "fSim": "apply_fsim", "GeneralizedfSim": "apply_fsim",
    }


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device path")
def test_backend_without_device_raises_import_error():
    from qibojit_amd.backends.hip import HipBackend

    with pytest.raises(ImportError):
        HipBackend()
    avail = qibojit_amd.MetaBackend().list_available()
    assert avail["hip"] is False
    assert set(avail) == {"numba", "cupy", "cuquantum", "hip"}


def test_qubits_tensor_and_custom_matrix():
    from qibojit_amd.backends.hip import HipBackend
    from qibojit_amd._qibo import gates

    g = gates.Unitary(np.eye(4), 3, 1).controlled_by(0, 4)
    np.testing.assert_array_equal(HipBackend._create_qubits_tensor(g, 6), [1, 2, 4, 5])
    assert HipBackend._create_qubits_tensor(g, 6).dtype == np.int32


def test_unknown_platform():
    with pytest.raises(ValueError):
        qibojit_amd.MetaBackend.load("tpu")


def test_scalar_angle_matrices_match_array_forms(dtype):
    """RX, RY, RZ, U1, U3 built from scalar math (double or integer angles)
    against the same formulas in NumPy, taken for angles the scalar forms
    do not accept (0-d arrays): equal to within an ulp of each precision,
    and float32 angles keep NumPy's single-precision evaluation."""
    from qibojit_amd.qibo_compat.npmatrices import NumpyMatrices, _scalars

    nm = NumpyMatrices(dtype)
    tol = 1e-15 if dtype == "complex128" else 2e-7
    rng = np.random.default_rng(31)
    for _ in range(200):
        a, b, c = rng.uniform(-8, 8, 3)
        assert _scalars(a, b, c) and not _scalars(np.array(a))
        for name, args in (("RX", (a,)), ("RY", (a,)), ("RZ", (a,)), ("U1", (a,)), ("U3", (a, b, c))):
            fast = getattr(nm, name)(*args)
            slow = getattr(nm, name)(*(np.array(x) for x in args))
            assert fast.dtype == slow.dtype == np.dtype(dtype)
            assert np.max(np.abs(fast - slow)) <= tol, (name, args)
    assert not _scalars(np.float32(0.3)) and not _scalars(True)
    assert nm.RY(3).dtype == np.dtype(dtype)
