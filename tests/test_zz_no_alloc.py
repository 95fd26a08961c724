"""qj_sample_shots_noalloc against qj_sample_shots, through the C-ABI.

The no-allocation entry point has no counterpart in the reference; it runs
the same CDF pipeline and search as qj_sample_shots (NumPy's
RandomState.choice, which CupyBackend.sample_shots defers to,
src/qibojit/backends/gpu.py:770-773) on a fixed per-GPU device array.  Its
samples must equal the allocating entry point's bit for bit, on consecutive
calls where a later call is smaller and so leaves stale data of the earlier
one in the fixed array.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_STATES = 1 << 20


def _run(name, probs, uniforms, nshots):
    import torch

    from qibojit_amd import _native

    code = _native.QJ_FLOAT64 if probs.dtype == torch.float64 else _native.QJ_FLOAT32
    out = torch.full((nshots,), -1, dtype=torch.int64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _native.call(name, ctypes.c_void_p(probs.data_ptr()), code, probs.numel(),
                 ctypes.c_void_p(uniforms.data_ptr()), nshots, ctypes.c_void_p(out.data_ptr()), stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _inputs(nstates, nshots, seed, realtype):
    import torch

    rng = np.random.default_rng(seed)
    p = rng.random(nstates) ** 3
    p[rng.integers(nstates, size=max(1, nstates // 7))] = 0
    if p.sum() == 0:
        p[0] = 1
    p = (p / p.sum()).astype(realtype)
    u = np.random.RandomState(seed).random_sample(nshots)
    return p, u, torch.from_numpy(p).cuda(), torch.from_numpy(u).cuda()


@pytest.mark.parametrize("realtype", [np.float64, np.float32])
def test_noalloc_matches_sample_shots_on_consecutive_calls(realtype):
    """Three calls in a row, each smaller than the one before: the largest
    allowed size (blocked scan over 64 block sums), a blocked scan over a
    few blocks, then the sequential cumsum.  Every call equals
    qj_sample_shots and NumPy's choice() on the same uniforms."""
    print(f"noalloc sample_shots, {np.dtype(realtype).name}", flush=True)
    for nstates, nshots, seed in ((MAX_STATES, 20000, 1), (2**17 + 3, 7000, 2), (1000, 3000, 3)):
        p, u, dp, du = _inputs(nstates, nshots, seed, realtype)
        got = _run("qj_sample_shots_noalloc", dp, du, nshots)
        want = _run("qj_sample_shots", dp, du, nshots)
        np.testing.assert_array_equal(got, want)
        ref = np.random.RandomState(seed).choice(nstates, size=nshots, p=p.astype(np.float64))
        np.testing.assert_array_equal(got, ref)
        assert np.all(p[got] > 0)


def test_noalloc_rejects_more_states_than_its_buffer():
    """More than 2^20 states is a ValueError raised before any launch; the
    output keeps its contents."""
    import torch

    from qibojit_amd import _native

    probs = torch.full((MAX_STATES + 1,), 1.0 / (MAX_STATES + 1), dtype=torch.float64, device="cuda")
    u = torch.zeros(4, dtype=torch.float64, device="cuda")
    out = torch.zeros(4, dtype=torch.int64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(ValueError, match="at most 1048576 states"):
        _native.call("qj_sample_shots_noalloc", ctypes.c_void_p(probs.data_ptr()), _native.QJ_FLOAT64,
                     MAX_STATES + 1, ctypes.c_void_p(u.data_ptr()), 4, ctypes.c_void_p(out.data_ptr()), stream)
    torch.cuda.synchronize()
    assert out.cpu().numpy().tolist() == [0, 0, 0, 0]
