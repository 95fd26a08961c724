"""Hardware-efficient ansatz on CPU: the oracle's pins and the host-side
layout of the batched engine (no GPU compute calls)."""

import ctypes

import numpy as np
import pytest

from oracle import ansatz_ref, einsum_ref
from oracle.gate_matrices import RY, RZ, X, random_state
from qibojit_amd import _native
from qibojit_amd.ansatz import (
    AnsatzEngine,
    HardwareEfficientAnsatz,
    ansatz_gradient_order,
    interleaved_groups,
    ry_matrices,
    rz_matrices,
)


def test_energy_of_zero_angles_is_the_ferromagnet():
    # RY(0) = RZ(0) = I and the ladder fixes |0...0>: E = -(n-1)
    for n in (2, 3, 5):
        assert ansatz_ref.hea_energy(n, 2, np.zeros(4 * n)) == pytest.approx(-(n - 1), abs=1e-14)


def test_parameter_shift_matches_finite_differences(rng):
    n, layers = 3, 2
    p = rng.uniform(-np.pi, np.pi, 2 * n * layers)
    ps = ansatz_ref.parameter_shift(n, layers, p, g=0.7)
    fd = ansatz_ref.central_difference(n, layers, p, g=0.7, h=1e-5)
    np.testing.assert_allclose(ps, fd, atol=1e-8)
    # a planted slip: shifting by pi/4 instead of pi/2 is caught
    wrong = []
    for i in range(p.size):
        a, b = p.copy(), p.copy()
        a[i] += np.pi / 4
        b[i] -= np.pi / 4
        wrong.append(0.5 * (ansatz_ref.hea_energy(n, layers, a, 0.7) - ansatz_ref.hea_energy(n, layers, b, 0.7)))
    assert np.max(np.abs(np.array(wrong) - fd)) > 1e-3


def test_single_qubit_pair_closed_form():
    # n = 2, one layer, only RY on qubit 0: |psi> = CNOT (c|0> + s|1>)|0> = c|00> + s|11>,
    # <ZZ> = 1 and <X_0> = <X_1> = 0, so E = -1 for every angle.
    for t in (0.3, 1.1, -2.0):
        p = np.zeros(4)
        p[0] = t
        assert ansatz_ref.hea_energy(2, 1, p) == pytest.approx(-1.0, abs=1e-14)
    # only RY on qubit 1: |0>(c|0> + s|1>) -> <ZZ> = cos t, <X_1> = sin t
    for t in (0.3, 1.1, -2.0):
        p = np.zeros(4)
        p[1] = t
        assert ansatz_ref.hea_energy(2, 1, p, g=0.5) == pytest.approx(-np.cos(t) - 0.5 * np.sin(t), abs=1e-14)


def test_ladder_permutation_is_the_cnot_ladder(rng):
    n = 5
    psi = random_state(n, rng)
    want = psi.copy()
    for q in range(n - 1):
        want = einsum_ref.apply_gate(want, X, [q + 1], controls=[q], nqubits=n)
    np.testing.assert_allclose(psi[ansatz_ref.ladder_permutation(n)], want, atol=1e-15)


def test_oracle_state_matches_the_port_backend(port_backend, rng):
    """The ansatz oracle agrees with the NumbaBackend restatement running
    the same gate list."""
    n, layers = 4, 2
    ansatz = HardwareEfficientAnsatz(n, layers)
    p = ansatz.random_parameters(rng)
    state = np.zeros(2**n, dtype=np.complex128)
    state[0] = 1
    for gate in ansatz.circuit(p).queue:
        state = port_backend.apply_gate(gate, state, n)
    np.testing.assert_allclose(state, ansatz_ref.hea_state(n, layers, p), atol=1e-13)


def test_rotation_matrices_match_textbook():
    for t in (-2.5, 0.0, 0.4, 3.0):
        np.testing.assert_allclose(ry_matrices(t), RY(t), atol=1e-15)
        np.testing.assert_allclose(rz_matrices(t), RZ(t), atol=1e-15)


def test_parameter_layout():
    a = HardwareEfficientAnsatz(5, 3)
    assert a.nparams == 30
    order = ansatz_gradient_order(a)
    assert sorted(order.tolist()) == list(range(30))
    gates_ = [g for g in a.circuit(np.arange(30.0)).queue if g.__class__.__name__ in ("RY", "RZ")]
    assert [g.parameters[0] for g in gates_] == [float(i) for i in order]
    with pytest.raises(ValueError):
        HardwareEfficientAnsatz(1, 2)


@pytest.mark.parametrize("n,k", [(24, 4), (24, 5), (10, 3), (7, 5), (3, 4)])
def test_interleaved_groups_partition_the_bits(n, k):
    groups = interleaved_groups(n, k)
    flat = sorted(b for grp in groups for b in grp)
    assert flat == list(range(n))
    assert all(len(grp) <= k and list(grp) == sorted(grp) for grp in groups)
    assert len(groups) == -(-n // k)
    # at most one of the lowest len(groups) bits per group
    assert all(sum(b < len(groups) for b in grp) == 1 for grp in groups)


def test_engine_packs_matrices_per_group_bit(rng):
    ansatz = HardwareEfficientAnsatz(7, 2)
    eng = AnsatzEngine(ansatz, 3, dtype="complex128", k_forward=3, k_adjoint=2, device="cpu")
    p = ansatz.random_parameters(rng, 3)
    fwd = eng._forward_mats(p).numpy()
    adj = eng._adjoint_mats(p).numpy()
    ang = p.reshape(3, 2, 2, 7)
    for layer in range(2):
        for i, bits in enumerate(eng.groups_f):
            for a, bit in enumerate(bits):
                q = 6 - bit
                for b in range(3):
                    want = RZ(ang[b, layer, 1, q]) @ RY(ang[b, layer, 0, q])
                    np.testing.assert_allclose(fwd[layer, i, b, a, 0].reshape(2, 2), want, atol=1e-15)
        for i, bits in enumerate(eng.groups_a):
            for a, bit in enumerate(bits):
                q = 6 - bit
                np.testing.assert_allclose(adj[layer, i, 1, a, 0].reshape(2, 2), RY(ang[1, layer, 0, q]), atol=1e-15)
                np.testing.assert_allclose(adj[layer, i, 1, a, 1].reshape(2, 2), RZ(ang[1, layer, 1, q]), atol=1e-15)


def test_engine_scatters_slot_sums_to_parameters():
    import torch

    ansatz = HardwareEfficientAnsatz(5, 2)
    eng = AnsatzEngine(ansatz, 2, dtype="complex64", k_adjoint=2, device="cpu")
    nl, ga, ka = 2, len(eng.groups_a), eng.k_adjoint
    sums = torch.full((nl, ga, ka * 2, 2), np.nan, dtype=torch.float64)
    for layer in range(nl):
        for i, bits in enumerate(eng.groups_a):
            for a, bit in enumerate(bits):
                for j in range(2):
                    for b in range(2):
                        sums[layer, i, a * 2 + j, b] = 1000 * b + ansatz.index(layer, j, 4 - bit)
    grads = eng._scatter(sums).numpy()
    want = np.arange(ansatz.nparams)[None, :] + 1000 * np.arange(2)[:, None]
    np.testing.assert_array_equal(grads, want)


def _fake():
    return ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first


def _rot(**kw):
    args = dict(psi_in=_fake(), psi_out=ctypes.c_void_p(0x2000), lam_in=None, lam_out=None,
                dtype=_native.QJ_COMPLEX64, n=10, stride=1 << 10, sel=None, nrows=4, bits=[0, 3, 6], nrot=1,
                gens=None, load=0, mats=_fake(), mstride=12, partials=None, ncols=4)
    args.update(kw)
    bits_ptr, _keep, nbits = _native.i32_array(args["bits"])
    gens = None
    if args["gens"] is not None:
        gens, _gkeep, _ = _native.i32_array(args["gens"])
    _native.call("qj_rot_pass", args["psi_in"], args["psi_out"], args["lam_in"], args["lam_out"], args["dtype"],
                 args["n"], args["stride"], args["sel"], args["nrows"], bits_ptr, nbits, args["nrot"], gens,
                 args["load"], args["mats"], args["mstride"], args["partials"], args["ncols"], None)


@pytest.mark.parametrize(
    "kw,match",
    [
        (dict(n=0), "nqubits"),
        (dict(nrows=0), "number of states"),
        (dict(stride=100), "stride"),
        (dict(bits=[3, 0]), "ascending"),
        (dict(bits=[0, 10]), "out of range"),
        (dict(bits=[0, 1, 2, 3, 4, 5]), "group size"),
        (dict(nrot=4), "step count"),
        (dict(load=7), "load mode"),
        (dict(ncols=0), "column"),
        (dict(mstride=3), "matrix stride"),
        (dict(load=1, psi_out=_fake()), "distinct"),
        (dict(lam_in=_fake()), "together"),
        (dict(lam_in=_fake(), lam_out=_fake(), bits=[0, 1, 2, 3, 4], mstride=20, dtype=_native.QJ_COMPLEX128),
         "at most 4"),
        (dict(lam_in=_fake(), lam_out=_fake(), nrot=2, mstride=24, gens=[2, 3, 2, 3, 2, 3]), "partials"),
        (dict(nrot=2, mstride=24, gens=[2, 3, 2, 9, 2, 3]), "generator"),
        (dict(dtype=_native.QJ_FLOAT32), "dtype"),
    ],
)
def test_rot_pass_argument_errors(kw, match):
    with pytest.raises(ValueError, match=match):
        _rot(**kw)


def test_tfim_and_reduce_argument_errors():
    with pytest.raises(ValueError, match="alias"):
        _native.call("qj_tfim_batch", _fake(), _fake(), _native.QJ_COMPLEX64, 6, 64, None, 1, 1.0, 0, _fake(), 1, None)
    with pytest.raises(ValueError, match="out of range"):
        _native.call("qj_tfim_batch", _fake(), None, _native.QJ_COMPLEX64, 1, 64, None, 1, 1.0, 0, _fake(), 1, None)
    with pytest.raises(ValueError, match="columns"):
        _native.call("qj_reduce_rows", _fake(), 3, 0, _fake(), None)


@pytest.mark.parametrize("n", [12, 13, 19, 20, 24, 30])
def test_tile_passes_cover_every_bit_once(n):
    from qibojit_amd.ansatz import TILE_BITS, TILE_LOW, tile_passes

    passes = tile_passes(n)
    rotated = [b for ps in passes for b in ps.rot]
    assert sorted(rotated) == list(range(n))
    for ps in passes:
        assert len(ps.bits) == TILE_BITS and list(ps.bits) == sorted(set(ps.bits))
        assert set(ps.rot) <= set(ps.bits) and max(ps.bits) < n
        assert set(range(TILE_LOW)) <= set(ps.bits)  # whole-line loads
    assert len(passes) == 1 + -(-(n - TILE_BITS) // (TILE_BITS - TILE_LOW))  # no pass wasted on wider runs


def test_tile_engine_packs_and_scatters(rng):
    import torch

    ansatz = HardwareEfficientAnsatz(14, 2)
    eng = AnsatzEngine(ansatz, 2, dtype="complex64", device="cpu", tiles_adjoint=True)
    assert eng.tiles and all(ps.kind == "tile" for ps in eng.passes_f + eng.passes_a)
    p = ansatz.random_parameters(rng, 2)
    adj = eng._adjoint_mats(p).numpy()
    ang = p.reshape(2, 2, 2, 14)
    for layer in range(2):
        for i, ps in enumerate(eng.passes_a):
            for a, bit in enumerate(ps.bits):
                m = adj[layer, i, 1, a]
                if bit in ps.rot:
                    np.testing.assert_allclose(m[0].reshape(2, 2), RY(ang[1, layer, 0, 13 - bit]), atol=1e-7)
                    np.testing.assert_allclose(m[1].reshape(2, 2), RZ(ang[1, layer, 1, 13 - bit]), atol=1e-7)
                else:
                    assert not m.any()
    sums = torch.full((2, len(eng.passes_a), eng.width_a * 2, 2), np.nan, dtype=torch.float64)
    for layer in range(2):
        for i, ps in enumerate(eng.passes_a):
            for a, bit in enumerate(ps.bits):
                if bit in ps.rot:
                    for j in range(2):
                        for b in range(2):
                            sums[layer, i, a * 2 + j, b] = 1000 * b + ansatz.index(layer, j, 13 - bit)
    grads = eng._scatter(sums).numpy()
    np.testing.assert_array_equal(grads, np.arange(ansatz.nparams)[None, :] + 1000 * np.arange(2)[:, None])
    with pytest.raises(ValueError, match="complex64"):
        AnsatzEngine(ansatz, 2, dtype="complex128", device="cpu", tiles=True)
