"""Host-side layout of the many-circuit ansatz runner (CPU, no GPU calls)."""

import numpy as np
import pytest
import torch

from oracle.gate_matrices import RY, RZ
from qibojit_amd.batch import LDS_MAX_QUBITS, _fused_layers, default_plan


def test_fused_layers_are_bit_indexed(rng):
    n, layers = 5, 3
    p = rng.uniform(-np.pi, np.pi, 2 * n * layers)
    m = _fused_layers(n, layers, p)
    ang = p.reshape(layers, 2, n)
    for layer in range(layers):
        for q in range(n):
            want = RZ(ang[layer, 1, q]) @ RY(ang[layer, 0, q])
            np.testing.assert_allclose(m[layer, n - 1 - q].reshape(2, 2), want, atol=1e-15)


def test_default_plan_uses_lds_up_to_its_limit():
    plan = default_plan(range(10, 23), torch.complex64)
    assert all(plan[n] == ("lds",) for n in range(10, LDS_MAX_QUBITS[torch.complex64] + 1))
    assert all(plan[n] == ("pass", 5) for n in range(LDS_MAX_QUBITS[torch.complex64] + 1, 23))
    plan = default_plan([11, 12, 13], torch.complex128)
    assert plan[13] == ("pass", 5) and plan[12] == ("lds",)


def test_lds_entry_rejects_oversized_circuits():
    import ctypes

    from qibojit_amd import _native

    fake = ctypes.c_void_p(0x1000)
    with pytest.raises(ValueError, match="LDS"):
        _native.call("qj_ansatz_lds", _native.QJ_COMPLEX64, fake, 4, fake, 100, 1.0, fake, 14, 4, None)
    with pytest.raises(ValueError, match="LDS"):
        _native.call("qj_ansatz_lds", _native.QJ_COMPLEX128, fake, 4, fake, 100, 1.0, fake, 13, 4, None)
    with pytest.raises(ValueError, match="workgroup"):
        _native.call("qj_ansatz_lds", _native.QJ_COMPLEX64, fake, 4, fake, 100, 1.0, fake, 10, 0, None)
