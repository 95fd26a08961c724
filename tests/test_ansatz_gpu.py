"""Batched ansatz engine, rotation passes, TFIM kernels and adjoint gradients
on the GPU, against the numpy oracle (parameter shift, finite differences,
closed forms)."""

import ctypes

import numpy as np
import pytest

from oracle import ansatz_ref, einsum_ref
from oracle.gate_matrices import RX as rx_matrix
from oracle.gate_matrices import RY as ry_matrix
from oracle.gate_matrices import RZ as rz_matrix
from oracle.gate_matrices import haar_unitary, random_state

pytestmark = pytest.mark.gpu

# amplitude / energy / gradient tolerances per precision
AMP = {"complex64": 1e-5, "complex128": 1e-12}
ENERGY = {"complex64": 2e-5, "complex128": 1e-11}
GRAD = {"complex64": 2e-5, "complex128": 1e-10}


@pytest.fixture(scope="module")
def torch_mod(hip_backend):  # the backend fixture proves the device and library
    import torch

    return torch


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("n,layers,batch,kf", [(6, 2, 3, 5), (7, 3, 2, 3), (9, 2, 4, 4), (5, 1, 1, 1), (10, 2, 2, 2)])
def test_engine_states_and_energies(torch_mod, dtype, rng, n, layers, batch, kf):
    from qibojit_amd.ansatz import AnsatzEngine, HardwareEfficientAnsatz

    ansatz = HardwareEfficientAnsatz(n, layers)
    eng = AnsatzEngine(ansatz, batch, dtype=dtype, g=0.8, k_forward=kf)
    p = ansatz.random_parameters(rng, batch)
    energies = _np(eng.energies(p))
    states = _np(eng.final_states())
    for b in range(batch):
        want = ansatz_ref.hea_state(n, layers, p[b])
        assert np.max(np.abs(states[b] - want)) < AMP[dtype]
        assert abs(energies[b] - einsum_ref.tfim_energy(want, 0.8)) < ENERGY[dtype]


@pytest.mark.parametrize("n,layers,batch,ka", [(6, 2, 2, 4), (7, 2, 3, 3), (5, 3, 1, 2), (8, 1, 2, 1), (11, 2, 2, 5)])
def test_engine_gradients_match_parameter_shift(torch_mod, dtype, rng, n, layers, batch, ka):
    from qibojit_amd.ansatz import AnsatzEngine, HardwareEfficientAnsatz

    ansatz = HardwareEfficientAnsatz(n, layers)
    eng = AnsatzEngine(ansatz, batch, dtype=dtype, g=1.0, k_adjoint=ka)
    p = ansatz.random_parameters(rng, batch)
    energies, grads = eng.gradients(p)
    energies, grads = _np(energies), _np(grads)
    for b in range(batch):
        ps = ansatz_ref.parameter_shift(n, layers, p[b], 1.0)
        assert np.max(np.abs(grads[b] - ps)) < GRAD[dtype], (b, np.max(np.abs(grads[b] - ps)))
        assert abs(energies[b] - ansatz_ref.hea_energy(n, layers, p[b], 1.0)) < ENERGY[dtype]
    # walking back through the circuit returns psi to |0...0>
    psi0 = _np(eng.psi[eng.cur])
    want = np.zeros_like(psi0)
    want[:, 0] = 1
    assert np.max(np.abs(psi0 - want)) < 10 * AMP[dtype]


def test_gradients_after_forward_reuse_the_states(torch_mod, rng):
    from qibojit_amd.ansatz import AnsatzEngine, HardwareEfficientAnsatz

    ansatz = HardwareEfficientAnsatz(6, 2)
    eng = AnsatzEngine(ansatz, 2, dtype="complex128")
    p = ansatz.random_parameters(rng, 2)
    e1 = _np(eng.energies(p))
    e2, g2 = eng.gradients(p)
    np.testing.assert_allclose(_np(e2), e1, atol=1e-13)
    # a second gradient call (forward states consumed) recomputes the forward
    e3, g3 = eng.gradients(p)
    np.testing.assert_allclose(_np(g3), _np(g2), atol=1e-13)
    np.testing.assert_allclose(_np(e3), e1, atol=1e-13)


def test_gradients_match_finite_differences(torch_mod, rng):
    from qibojit_amd.ansatz import AnsatzEngine, HardwareEfficientAnsatz

    n, layers = 6, 2
    ansatz = HardwareEfficientAnsatz(n, layers)
    p = ansatz.random_parameters(rng, 1)
    _, grads = AnsatzEngine(ansatz, 1, dtype="complex128").gradients(p)
    # central differences in complex128 through the engine itself, h = 1e-4
    h = 1e-4
    shifted = np.repeat(p, 2 * ansatz.nparams, axis=0)
    for i in range(ansatz.nparams):
        shifted[2 * i, i] += h
        shifted[2 * i + 1, i] -= h
    e = _np(AnsatzEngine(ansatz, shifted.shape[0], dtype="complex128").energies(shifted))
    fd = (e[0::2] - e[1::2]) / (2 * h)
    np.testing.assert_allclose(_np(grads)[0], fd, atol=1e-7)


def test_plain_adjoint_matches_parameter_shift(backend, rng):
    from qibojit_amd.ansatz import HardwareEfficientAnsatz, ansatz_gradient_order, plain_adjoint

    n, layers = 6, 2
    ansatz = HardwareEfficientAnsatz(n, layers)
    p = ansatz.random_parameters(rng)
    energy, grads = plain_adjoint(backend, ansatz.circuit(p), g=1.0)
    ps = ansatz_ref.parameter_shift(n, layers, p, 1.0)
    np.testing.assert_allclose(grads, ps[ansatz_gradient_order(ansatz)], atol=1e-10)
    assert energy == pytest.approx(ansatz_ref.hea_energy(n, layers, p, 1.0), abs=1e-11)


def test_autograd_function(torch_mod, rng):
    from qibojit_amd.ansatz import AnsatzEnergy, AnsatzEngine, HardwareEfficientAnsatz

    torch = torch_mod
    n, layers, batch = 5, 2, 3
    ansatz = HardwareEfficientAnsatz(n, layers)
    eng = AnsatzEngine(ansatz, batch, dtype="complex128")
    p = torch.tensor(ansatz.random_parameters(rng, batch), dtype=torch.float64, requires_grad=True)
    weights = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    loss = (AnsatzEnergy.apply(p, eng).cpu() * weights).sum()
    loss.backward()
    pn = p.detach().numpy()
    for b in range(batch):
        ps = ansatz_ref.parameter_shift(n, layers, pn[b], 1.0)
        np.testing.assert_allclose(p.grad[b].numpy(), weights[b].item() * ps, atol=1e-10)


def _rot_call(psi_in, psi_out, lam_in, lam_out, code, n, stride, sel, nrows, bits, nrot, gens, load, mats, mstride,
              partials, ncols):
    from qibojit_amd import _native

    bits_ptr, _keep, nbits = _native.i32_array(bits)
    gptr = None
    if gens is not None:
        gptr, _gkeep, _ = _native.i32_array(gens)

    def vp(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else None

    _native.call("qj_rot_pass", vp(psi_in), vp(psi_out), vp(lam_in), vp(lam_out), code, n, stride, vp(sel), nrows,
                 bits_ptr, nbits, nrot, gptr, load, vp(mats), mstride, vp(partials), ncols, None)


@pytest.mark.parametrize("load", [0, 1, 2])
@pytest.mark.parametrize("bits", [(0, 4, 7), (1, 2), (3,), (0, 1, 2, 3, 4), (2, 5, 6, 8)])
def test_rot_pass_forward_against_numpy(torch_mod, dtype, rng, load, bits):
    """Every load mode, with and without bit 0, several steps per bit, on a
    selected subset of a batch (unselected states stay untouched)."""
    from qibojit_amd import _native

    torch = torch_mod
    n, nstates, nrot = 9, 4, 2
    code = _native.QJ_COMPLEX64 if dtype == "complex64" else _native.QJ_COMPLEX128
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    host = np.stack([random_state(n, rng) for _ in range(nstates)]).astype(dtype)
    k = len(bits)
    mats = np.stack([np.stack([[haar_unitary(2, rng) for _ in range(nrot)] for _ in range(k)]) for _ in range(nstates)])
    sel = [3, 1]
    src = torch.from_numpy(host).cuda()
    dst = src.clone() if load == 0 else torch.zeros_like(src)
    _rot_call(src, dst, None, None, code, n, 1 << n, torch.tensor(sel, dtype=torch.int32).cuda(), len(sel),
              list(bits), nrot, None, load, torch.from_numpy(mats.astype(dtype).reshape(nstates, -1)).cuda(),
              k * nrot * 4, None, 3)
    got = _np(dst)
    idx = np.arange(1 << n)
    for s in range(nstates):
        if s not in sel:
            np.testing.assert_array_equal(got[s], host[s] if load == 0 else 0)
            continue
        x = host[s].astype(np.complex128)
        if load == 1:
            x = x[idx ^ (idx >> 1)]
        elif load == 2:
            x = x[np.argsort(idx ^ (idx >> 1))]
        for a, bit in enumerate(bits):
            for j in range(nrot):
                x = einsum_ref.apply_gate(x, mats[s, a, j], [n - 1 - bit], nqubits=n)
        assert np.max(np.abs(got[s] - x)) < AMP[dtype]


@pytest.mark.parametrize("bits,forms", [((0, 3, 5), (1, 2, 3)), ((1, 2, 4, 6), (2, 0, 3)), ((0, 1), (3, 1, 0))])
def test_rot_pass_generator_forms(torch_mod, dtype, rng, bits, forms):
    """Steps declared by their generator (X, Y, Z) run in the X/Y/Z matrix
    forms, forward and adjoint (psi, lam and Im<lam|G|psi>), against numpy;
    a step without a generator keeps the general 2x2 path."""
    torch = torch_mod
    from qibojit_amd import _native

    n, nstates = 8, 2
    code = _native.QJ_COMPLEX64 if dtype == "complex64" else _native.QJ_COMPLEX128
    paulis = {1: np.array([[0, 1], [1, 0]]), 2: np.array([[0, -1j], [1j, 0]]), 3: np.diag([1.0, -1.0])}
    k, nrot = len(bits), len(forms)
    mats = np.zeros((nstates, k, nrot, 2, 2), dtype=np.complex128)
    for s in range(nstates):
        for a in range(k):
            th = rng.uniform(-np.pi, np.pi, size=nrot)
            by_form = {1: rx_matrix, 2: ry_matrix, 3: rz_matrix}
            mats[s, a] = [by_form[f](th[j]) if f else haar_unitary(2, rng) for j, f in enumerate(forms)]
    gens = [g for _ in bits for g in forms]
    dm = torch.from_numpy(mats.astype(dtype).reshape(nstates, -1)).cuda()
    host = np.stack([random_state(n, rng) for _ in range(nstates)]).astype(dtype)
    lamh = np.stack([random_state(n, rng) for _ in range(nstates)]).astype(dtype)
    # forward with the forms declared
    src = torch.from_numpy(host).cuda()
    dst = src.clone()
    _rot_call(src, dst, None, None, code, n, 1 << n, None, nstates, list(bits), nrot, gens, 0, dm, k * nrot * 4,
              None, 1)
    for s in range(nstates):
        x = host[s].astype(np.complex128)
        for a, bit in enumerate(bits):
            for j in range(nrot):
                x = einsum_ref.apply_gate(x, mats[s, a, j], [n - 1 - bit], nqubits=n)
        assert np.max(np.abs(_np(dst)[s] - x)) < AMP[dtype]
    # adjoint
    psi, lam = torch.from_numpy(host).cuda(), torch.from_numpy(lamh).cuda()
    pout, lout = psi.clone(), lam.clone()
    ncols = 2
    part = torch.zeros((k * nrot, nstates, ncols), dtype=torch.float64).cuda()
    _rot_call(psi, pout, lam, lout, code, n, 1 << n, None, nstates, list(bits), nrot, gens, 0, dm, k * nrot * 4,
              part, ncols)
    grad = _np(part).sum(axis=-1)
    for s in range(nstates):
        x, y = host[s].astype(np.complex128), lamh[s].astype(np.complex128)
        for a, bit in enumerate(bits):
            t = [n - 1 - bit]
            for j in reversed(range(nrot)):
                if forms[j]:
                    gx = einsum_ref.apply_gate(x, paulis[forms[j]], t, nqubits=n)
                    assert abs(grad[a * nrot + j, s] - np.vdot(y, gx).imag) < 10 * AMP[dtype]
                u = mats[s, a, j].conj().T
                x = einsum_ref.apply_gate(x, u, t, nqubits=n)
                y = einsum_ref.apply_gate(y, u, t, nqubits=n)
        assert np.max(np.abs(_np(pout)[s] - x)) < AMP[dtype]
        assert np.max(np.abs(_np(lout)[s] - y)) < AMP[dtype]


def test_rot_pass_zero_load_builds_product_state(torch_mod, rng):
    from qibojit_amd import _native

    torch = torch_mod
    n, bits = 6, (0, 2, 5)
    mats = np.stack([haar_unitary(2, rng) for _ in bits])[None, :, None]  # (1, k, 1, 2, 2)
    dst = torch.full((1 << n,), 7.0 + 7.0j, dtype=torch.complex128).cuda()
    _rot_call(None, dst, None, None, _native.QJ_COMPLEX128, n, 1 << n, None, 1, list(bits), 1, None, 3,
              torch.from_numpy(mats.reshape(-1)).cuda(), 12, None, 1)
    x = np.zeros(1 << n, dtype=np.complex128)
    x[0] = 1
    for a, bit in enumerate(bits):
        x = einsum_ref.apply_gate(x, mats[0, a, 0], [n - 1 - bit], nqubits=n)
    np.testing.assert_allclose(_np(dst), x, atol=1e-14)


_TFIM_MATRICES = {}


# n = 11 with ncols 1 and 3: several grid-stride iterations per workgroup
# and unevenly loaded workgroups
@pytest.mark.parametrize("n,ncols", [(8, 2), (11, 1), (11, 3)])
@pytest.mark.parametrize("ladder", [0, 1])
def test_tfim_batch_energy_and_apply(torch_mod, dtype, rng, ladder, n, ncols):
    from qibojit_amd import _native

    torch = torch_mod
    nstates, g = 3, 0.6
    code = _native.QJ_COMPLEX64 if dtype == "complex64" else _native.QJ_COMPLEX128
    host = np.stack([random_state(n, rng) for _ in range(nstates)]).astype(dtype)
    psi = torch.from_numpy(host).cuda()
    out = torch.zeros_like(psi)
    part = torch.zeros((nstates, ncols), dtype=torch.float64).cuda()
    part2 = torch.zeros_like(part)
    sums = torch.zeros(nstates, dtype=torch.float64).cuda()
    sums2 = torch.zeros_like(sums)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _native.call("qj_tfim_batch", vp(psi), vp(out), code, n, 1 << n, None, nstates, g, ladder, vp(part), ncols, None)
    _native.call("qj_tfim_batch", vp(psi), None, code, n, 1 << n, None, nstates, g, ladder, vp(part2), ncols, None)
    _native.call("qj_reduce_rows", vp(part), nstates, ncols, vp(sums), None)
    _native.call("qj_reduce_rows", vp(part2), nstates, ncols, vp(sums2), None)
    if (n, g) not in _TFIM_MATRICES:
        _TFIM_MATRICES[(n, g)] = einsum_ref.tfim_matrix(n, g)
    h = _TFIM_MATRICES[(n, g)]
    perm = ansatz_ref.ladder_permutation(n)
    inv = np.argsort(perm)
    for s in range(nstates):
        x = host[s].astype(np.complex128)
        if ladder:
            hx = (h @ x[perm])[inv]  # L^+ H L x
        else:
            hx = h @ x
        e = np.vdot(x, hx).real
        tol = 50 * AMP[dtype]
        assert np.max(np.abs(_np(out)[s] - hx)) < tol
        assert abs(_np(sums)[s] - e) < ENERGY[dtype]
        assert abs(_np(sums2)[s] - e) < ENERGY[dtype]


@pytest.mark.parametrize("n,layers,batch", [(12, 2, 2), (14, 3, 3), (21, 1, 1)])
def test_tile_engine_energies_states_and_gradients(torch_mod, rng, n, layers, batch):
    """LDS tile passes (complex64, n >= 12) against the oracle and against
    the register passes."""
    from qibojit_amd.ansatz import AnsatzEngine, HardwareEfficientAnsatz

    ansatz = HardwareEfficientAnsatz(n, layers)
    p = ansatz.random_parameters(rng, batch)
    tile = AnsatzEngine(ansatz, batch, dtype="complex64", tiles=True, tiles_adjoint=True)
    reg = AnsatzEngine(ansatz, batch, dtype="complex64", tiles=False, tiles_adjoint=False)
    e_t, g_t = tile.gradients(p)
    e_r, g_r = reg.gradients(p)
    np.testing.assert_allclose(_np(e_t), _np(e_r), atol=ENERGY["complex64"])
    assert np.max(np.abs(_np(g_t) - _np(g_r))) < GRAD["complex64"]
    tile.energies(p)
    states = _np(tile.final_states())
    if n <= 14:
        for b in range(batch):
            want = ansatz_ref.hea_state(n, layers, p[b])
            assert np.max(np.abs(states[b] - want)) < AMP["complex64"]
        ps = ansatz_ref.parameter_shift(n, layers, p[0], 1.0, indices=range(0, ansatz.nparams, 5))
        assert np.max(np.abs(_np(g_t)[0][::5] - ps)) < GRAD["complex64"]


@pytest.mark.parametrize("load", [0, 1, 2])
def test_rot_tile_pass_against_numpy(torch_mod, rng, load):
    """Tile pass with a partial rotation mask, two steps, every load mode,
    forward and adjoint (gradient partials and both output states)."""
    from qibojit_amd import _native

    torch = torch_mod
    n, nstates, nrot = 14, 3, 2
    tile_bits = [0, 1, 2, 3, 5, 6, 8, 9, 10, 11, 12, 13]
    rot = [1, 3, 4, 6, 7, 9, 11]  # tile indices rotated
    rotmask = sum(1 << j for j in rot)
    host = np.stack([random_state(n, rng) for _ in range(nstates)]).astype(np.complex64)
    lamh = np.stack([random_state(n, rng) for _ in range(nstates)]).astype(np.complex64)
    # steps with a generator are exp(-i theta G / 2) (the form the kernel
    # relies on); the other steps of rotated bits are general unitaries
    mats = np.zeros((nstates, 12, nrot, 2, 2), dtype=np.complex128)
    for s in range(nstates):
        for j in range(12):
            for st in range(nrot):
                if j in rot:
                    th = rng.uniform(-np.pi, np.pi)
                    mats[s, j, st] = (ry_matrix(th), rz_matrix(th))[st]
                else:
                    mats[s, j, st] = haar_unitary(2, rng)
    gens = [g for j in range(12) for g in ((2, 3) if j in rot else (0, 0))]
    idx = np.arange(1 << n)

    def vp(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else None

    bits_ptr, _k1, _ = _native.i32_array(tile_bits)
    gptr, _k2, _ = _native.i32_array(gens)
    dm = torch.from_numpy(mats.astype(np.complex64).reshape(nstates, -1)).cuda()

    def permuted(x):
        if load == 1:
            return x[idx ^ (idx >> 1)]
        if load == 2:
            return x[np.argsort(idx ^ (idx >> 1))]
        return x

    # forward, without and with the generators (with them, the Z steps of
    # every rotated bit run as one diagonal after the RY steps)
    for gen_ptr in (None, gptr):
        src = torch.from_numpy(host).cuda()
        dst = src.clone() if load == 0 else torch.zeros_like(src)
        _native.call("qj_rot_tile_pass", vp(src), vp(dst), None, None, _native.QJ_COMPLEX64, n, 1 << n, None,
                     nstates, bits_ptr, rotmask, nrot, gen_ptr, load, vp(dm), 12 * nrot * 4, None, 2, None)
        got = _np(dst)
        for s in range(nstates):
            x = permuted(host[s].astype(np.complex128))
            for j in rot:
                for st in range(nrot):
                    x = einsum_ref.apply_gate(x, mats[s, j, st], [n - 1 - tile_bits[j]], nqubits=n)
            assert np.max(np.abs(got[s] - x)) < AMP["complex64"]
    # adjoint
    src = torch.from_numpy(host).cuda()
    lsrc = torch.from_numpy(lamh).cuda()
    dst = src.clone() if load == 0 else torch.zeros_like(src)
    ldst = lsrc.clone() if load == 0 else torch.zeros_like(lsrc)
    ncols = 3
    part = torch.zeros((12 * nrot, nstates, ncols), dtype=torch.float64).cuda()
    _native.call("qj_rot_tile_pass", vp(src), vp(dst), vp(lsrc), vp(ldst), _native.QJ_COMPLEX64, n, 1 << n, None,
                 nstates, bits_ptr, rotmask, nrot, gptr, load, vp(dm), 12 * nrot * 4, vp(part), ncols, None)
    got, lgot, grad = _np(dst), _np(ldst), _np(part).sum(axis=-1)
    paulis = {2: np.array([[0, -1j], [1j, 0]]), 3: np.diag([1.0, -1.0])}
    for s in range(nstates):
        x = permuted(host[s].astype(np.complex128))
        y = permuted(lamh[s].astype(np.complex128))
        for j in rot:
            t = [n - 1 - tile_bits[j]]
            for st in reversed(range(nrot)):
                gx = einsum_ref.apply_gate(x, paulis[(2, 3)[st]], t, nqubits=n)
                assert abs(grad[j * nrot + st, s] - np.vdot(y, gx).imag) < 1e-4
                u = mats[s, j, st].conj().T
                x = einsum_ref.apply_gate(x, u, t, nqubits=n)
                y = einsum_ref.apply_gate(y, u, t, nqubits=n)
        assert np.max(np.abs(got[s] - x)) < AMP["complex64"]
        assert np.max(np.abs(lgot[s] - y)) < AMP["complex64"]
