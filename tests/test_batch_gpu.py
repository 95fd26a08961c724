"""Many ansatz circuits of uneven shape: LDS and pass paths, graph capture
and replay, autotuner, against the numpy oracle."""

import numpy as np
import pytest

from oracle import ansatz_ref

pytestmark = pytest.mark.gpu

ENERGY = {"complex64": 2e-5, "complex128": 1e-11}


def _shapes_and_params(rng, shapes):
    params = [rng.uniform(-np.pi, np.pi, 2 * n * nl) for n, nl in shapes]
    want = np.array([ansatz_ref.hea_energy(n, nl, p, 1.0) for (n, nl), p in zip(shapes, params)])
    return params, want


SHAPES = [(3, 2), (9, 1), (5, 3), (14, 2), (12, 4), (2, 1), (15, 1), (9, 3), (14, 1), (13, 2), (4, 5)]


@pytest.mark.parametrize("mode", ["default", "all_pass_k2", "all_pass_k5", "mixed", "tiles"])
def test_batch_energies(hip_backend, dtype, rng, mode):
    from qibojit_amd.batch import AnsatzBatch

    params, want = _shapes_and_params(rng, SHAPES)
    ns = sorted({n for n, _ in SHAPES})
    if mode == "tiles" and dtype != "complex64":
        pytest.skip("tile passes are complex64")
    if mode == "default":
        plan = None
    elif mode == "tiles":
        plan = {n: ("tile",) if n >= 12 else ("lds",) for n in ns}
    elif mode.startswith("all_pass"):
        k = int(mode[-1])
        plan = {n: ("pass", min(k, n)) for n in ns}
    else:
        plan = {n: (("lds",) if n % 2 and n <= 12 else ("pass", 3 if n > 2 else 2)) for n in ns}
    b = AnsatzBatch(SHAPES, dtype=dtype, plan=plan)
    b.set_parameters(params)
    b.run()
    got = b.energies().cpu().numpy()
    assert np.max(np.abs(got - want)) < ENERGY[dtype]


def test_graph_capture_and_replay_with_new_parameters(hip_backend, rng):
    import torch

    from qibojit_amd.batch import AnsatzBatch

    params, want = _shapes_and_params(rng, SHAPES)
    b = AnsatzBatch(SHAPES, dtype="complex128")
    b.set_parameters(params)
    b.capture()
    for rep in range(3):
        b.energies_sorted.zero_()
        b.replay()
        torch.cuda.synchronize()
        got = b.energies().cpu().numpy()
        assert np.max(np.abs(got - want)) < ENERGY["complex128"], rep
        # eager work between replays
        b.run()
        torch.cuda.synchronize()
    params2, want2 = _shapes_and_params(rng, SHAPES)
    b.set_parameters(params2)
    b.replay()
    torch.cuda.synchronize()
    assert np.max(np.abs(b.energies().cpu().numpy() - want2)) < ENERGY["complex128"]


def test_autotune_plan_is_valid_and_correct(hip_backend, rng):
    from qibojit_amd.batch import LDS_MAX_QUBITS, AnsatzBatch, autotune

    shapes = [(4, 2), (14, 1), (4, 1), (13, 2), (14, 2)]
    params, want = _shapes_and_params(rng, shapes)
    plan, timings = autotune(shapes, params, reps=1)
    assert set(plan) == {4, 13, 14}
    for n, choice in plan.items():
        assert choice in timings[n] and timings[n][choice] == min(timings[n].values())
        if choice[0] == "lds":
            import torch

            assert n <= LDS_MAX_QUBITS[torch.complex64]
    b = AnsatzBatch(shapes, plan=plan)
    b.set_parameters(params)
    b.run()
    assert np.max(np.abs(b.energies().cpu().numpy() - want)) < ENERGY["complex64"]


def test_plain_energies_match_oracle(backend, rng):
    from qibojit_amd.batch import plain_energies

    shapes = [(5, 2), (7, 1), (3, 3)]
    params, want = _shapes_and_params(rng, shapes)
    got = plain_energies(backend, shapes, params)
    np.testing.assert_allclose(got, want, atol=1e-11)


def test_lds_kernel_marks_bad_descriptors_nan(hip_backend):
    import ctypes

    import torch

    from qibojit_amd import _native

    desc = torch.tensor([[3, 1, 0, 0], [14, 1, 0, 0], [3, 0, 0, 0], [3, 2, 1000, 0]], dtype=torch.int32).cuda()
    mats = torch.zeros(24, dtype=torch.complex64).cuda()
    mats[0::4] = 1
    mats[3::4] = 1  # identity rotations
    out = torch.zeros(4, dtype=torch.float64).cuda()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _native.call("qj_ansatz_lds", _native.QJ_COMPLEX64, vp(desc), 4, vp(mats), mats.numel(), 1.0, vp(out), 13, 2, None)
    got = out.cpu().numpy()
    assert got[0] == pytest.approx(-2.0)  # |000>: E = -(n-1)
    assert np.isnan(got[1:]).all()
