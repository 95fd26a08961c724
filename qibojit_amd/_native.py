"""ctypes binding of ``libqjhip.so`` (declared in include/qjhip.h).

The library is loaded after ``torch`` so that both share one HIP runtime:
torch's bundled ``libamdhip64.so`` carries the soname ``libamdhip64.so.7``
that the library needs, so the dynamic loader reuses it instead of mapping a
second runtime.  Device memory and streams are torch's; every computation is
a kernel of this library.

There is no fallback: if the library is missing, :func:`lib` raises.
"""

import ctypes
import os

import torch  # noqa: F401  (must precede the library load, see module doc)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqjhip.so")
ABI_VERSION = 1

QJ_OK, QJ_EVALUE, QJ_EMEMORY, QJ_EHIP = 0, 1, 2, 3
QJ_COMPLEX64, QJ_COMPLEX128, QJ_FLOAT32, QJ_FLOAT64 = 0, 1, 2, 3

# kernel kinds, keyed by the reference's kernel names (cpu.py:23-36)
ONE_QUBIT_KINDS = {"apply_gate": 0, "apply_x": 1, "apply_y": 2, "apply_z": 3, "apply_z_pow": 4}
TWO_QUBIT_KINDS = {"apply_two_qubit_gate": 0, "apply_swap": 1, "apply_fsim": 2}

_vp, _i32, _i64, _u64, _dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pu32 = ctypes.POINTER(ctypes.c_uint32)
_pdbl = ctypes.POINTER(ctypes.c_double)

# name -> argument types (all return int status, except qj_last_error)
SIGNATURES = {
    "qj_abi_version": [],
    "qj_device_count": [ctypes.POINTER(ctypes.c_int)],
    "qj_set_device": [_i32],
    "qj_malloc": [ctypes.POINTER(_vp), _u64],
    "qj_free": [_vp],
    "qj_memcpy_htod": [_vp, _vp, _u64, _vp],
    "qj_memcpy_dtoh": [_vp, _vp, _u64, _vp],
    "qj_memcpy_dtod": [_vp, _vp, _u64, _vp],
    "qj_stream_synchronize": [_vp],
    "qj_initial_state": [_vp, _i32, _u64, _vp],
    "qj_apply_one_qubit": [_vp, _i32, _i32, _i32, _i32, _pi32, _i32, _vp, _vp],
    "qj_apply_two_qubit": [_vp, _i32, _i32, _i32, _i32, _i32, _pi32, _i32, _vp, _vp],
    "qj_apply_multi_qubit": [_vp, _i32, _i32, _pi32, _i32, _pi32, _i32, _vp, _vp],
    "qj_apply_tile_pass": [_vp, _i32, _i32, _i32, _pi32, _i32, _vp, _i32, _vp, _u64, _vp],
    "qj_apply_bit_swaps": [_vp, _i32, _i32, _pi32, _i32, _vp],
    "qj_collapse_state": [_vp, _i32, _i32, _pi32, _i32, _i64, _i32, _pdbl, _vp],
    "qj_calculate_probabilities": [_vp, _i32, _i32, _pi32, _i32, _vp, _i32, _vp],
    "qj_calculate_probabilities_dm": [_vp, _i32, _i32, _pi32, _i32, _vp, _i32, _vp],
    "qj_sample_shots": [_vp, _i32, _i64, _vp, _i64, _vp, _vp],
    "qj_sample_shots_noalloc": [_vp, _i32, _i64, _vp, _i64, _vp, _vp],
    "qj_count_samples": [_vp, _i64, _vp, _i64, _vp, _vp],
    "qj_measure_frequencies": [_vp, _vp, _i32, _i64, _i64, _pu32, _i32, _vp],
    "qj_norm_squared": [_vp, _i32, _u64, _pdbl, _vp],
    "qj_vdot": [_vp, _vp, _i32, _u64, _pdbl, _vp],
    "qj_expectation_tfim": [_vp, _i32, _i32, _dbl, _pdbl, _vp],
    "qj_axpby": [_vp, _vp, _i32, _u64, _pdbl, _pdbl, _vp],
    "qj_rot_pass": [_vp, _vp, _vp, _vp, _i32, _i32, _u64, _vp, _i32, _pi32, _i32, _i32, _pi32, _i32,
                    _vp, _u64, _vp, _i32, _vp],
    "qj_rot_tile_pass": [_vp, _vp, _vp, _vp, _i32, _i32, _u64, _vp, _i32, _pi32, _i32, _i32, _pi32, _i32,
                         _vp, _u64, _vp, _i32, _vp],
    "qj_tfim_batch": [_vp, _vp, _i32, _i32, _u64, _vp, _i32, _dbl, _i32, _vp, _i32, _vp],
    "qj_reduce_rows": [_vp, _u64, _u64, _vp, _vp],
    "qj_fill_state": [_vp, _i32, _u64, _dbl, _dbl, _vp],
    "qj_compare_states": [_vp, _i32, _vp, _i32, _u64, _pdbl, _vp],
    "qj_apply_block_mfma": [_vp, _i32, _i32, _pi32, _i32, _pi32, _i32, _vp, _vp],
    "qj_ansatz_lds": [_i32, _vp, _i32, _vp, _u64, _dbl, _vp, _i32, _i32, _vp],
    "qj_clifford_pack": [_vp, _i32, _vp, _vp],
    "qj_clifford_unpack": [_vp, _i32, _vp, _vp],
    "qj_clifford_apply_gate": [_vp, _i32, _i32, _i32, _i32, _vp],
    "qj_clifford_apply_circuit": [_vp, _i32, _pi32, _i64, _vp],
    "qj_clifford_schedule": [_pi32, _i64, _i32, _vp, _i64, ctypes.POINTER(ctypes.c_int64)],
    "qj_clifford_apply_schedule": [_vp, _i32, _vp, _i64, _vp],
    "qj_clifford_first_stabilizer": [_vp, _i32, _i32, _vp, _vp],
    "qj_clifford_random_outcome": [_vp, _i32, _i32, _i32, _i32, _vp],
    "qj_clifford_determined_outcome": [_vp, _i32, _i32, _vp, _vp],
    "qj_clifford_determined_outcome_exact": [_vp, _i32, _i32, _vp, _vp],
    "qj_clifford_measure": [_vp, _i32, _vp, _i32, _vp, _i32, _vp, _vp, _vp],
    "qj_clifford_rowsum": [_vp, _i32, _vp, _vp, _i64, _i32, _vp],
    "qj_clifford_sym_init": [_vp, _i32, _vp, _i32, _vp],
    "qj_clifford_sym_random": [_vp, _i32, _vp, _i32, _i32, _i32, _i32, _vp],
    "qj_clifford_sym_determined": [_vp, _i32, _vp, _i32, _i32, _vp, _vp],
    "qj_clifford_sym_determined_exact": [_vp, _i32, _vp, _i32, _i32, _vp, _vp],
    "qj_gf2_samples": [_vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp],
    "qj_packbits_rows": [_vp, _i64, _i64, _vp, _vp],
    "qj_unpackbits_rows": [_vp, _i64, _i64, _i64, _vp, _vp],
}

# entry points that return a value instead of a status
VALUE_SIGNATURES = {
    "qj_clifford_column_words": (ctypes.c_int64, [_i32]),
}

_LIB = None


class NativeLibraryMissing(ImportError):
    """libqjhip.so has not been built (run __graft_entry__.build())."""


def lib():
    """Load (once) and return the configured ctypes library."""
    global _LIB  # pylint: disable=global-statement
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} not found: build the HIP library first (qibojit_amd.build.build())."
        )
    handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
    handle.qj_last_error.restype = ctypes.c_char_p
    handle.qj_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    for name, (res, args) in VALUE_SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    if handle.qj_abi_version() != ABI_VERSION:
        raise NativeLibraryMissing("libqjhip.so ABI version mismatch: rebuild the library.")
    _LIB = handle
    return _LIB


def check(status):
    """Map a status code to the Python exception the reference raises."""
    if status == QJ_OK:
        return
    msg = lib().qj_last_error().decode(errors="replace")
    if status == QJ_EVALUE:
        raise ValueError(msg)
    if status == QJ_EMEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def device_count():
    n = ctypes.c_int(0)
    call("qj_device_count", ctypes.byref(n))
    return n.value


def i32_array(values):
    """(ctypes pointer, keepalive, length) for a sequence of int32."""
    vals = [int(v) for v in values]
    arr = (ctypes.c_int32 * max(1, len(vals)))(*vals)
    return ctypes.cast(arr, _pi32), arr, len(vals)
