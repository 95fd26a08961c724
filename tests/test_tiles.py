"""Fused tile planner and encoder, on the CPU: plan + encode + a numpy
emulation of the documented pass semantics equals gate-by-gate execution
by the CPU port of NumbaBackend."""

import numpy as np
import pytest

from oracle import einsum_ref as er
from oracle import gate_matrices as gm
from oracle import port
from qibojit_amd import tiles
from qibojit_amd.qibo_compat import gates
from tests.tile_emulator import apply_bit_swaps, apply_tile_pass


def _run_planned(gate_list, psi, nbits, low, max_high, density_nq=None):
    """Same loop as HipBackend.apply_gates_fused, with the emulator."""
    out = psi.copy()
    pending, all_passes = [], []

    def flush(state):
        if pending:
            passes = tiles.plan(pending, nbits, low, max_high)
            for p in passes:
                if isinstance(p, tiles.SwapPass):
                    state = apply_bit_swaps(state, nbits, p.pairs)
                    continue
                assert len(p.high) <= max_high and all(b >= low for b in p.high)
                enc_ops, data, high = tiles.encode(p, nbits, psi.dtype)
                state = apply_tile_pass(state, nbits, p.low, high, enc_ops, data, psi.dtype)
            all_passes.extend(passes)
            pending.clear()
        return state

    for g in gate_list:
        for h in (tiles.expand_fanout(g) if g.name == "fanout" else [g]):
            op = tiles.gate_to_op(port.custom_matrix, h, nbits)
            if op is None or not tiles.fits_alone(op, low, max_high):
                out = flush(out)
                nq = density_nq or nbits
                shaped = out.reshape(2**nq, 2**nq) if density_nq else out
                out = port.PortBackend(2).apply_gate(h, shaped.copy(), nq).reshape(-1)
                continue
            pending.append(op)
            if density_nq is not None:
                pending.append(tiles.column_op(op, density_nq))
    out = flush(out)
    return out, all_passes


def _random_circuit(n, depth, rng):
    out = []
    for layer in range(depth):
        for q in range(n):
            out.append(gates.U3(q, *rng.uniform(0, 2 * np.pi, 3)))
            out.append(gates.RZ(q, rng.uniform(0, 2 * np.pi)))
        for q in range(layer % 2, n - 1, 2):
            out.append(gates.CZ(q, q + 1))
        a, b = rng.choice(n, 2, replace=False)
        out.append(gates.CNOT(int(a), int(b)))
        a, b = rng.choice(n, 2, replace=False)
        out.append(gates.fSim(int(a), int(b), 0.3, 0.7))
        a, b, c = rng.choice(n, 3, replace=False)
        out.append(gates.Unitary(gm.haar_unitary(8, rng), int(a), int(b), int(c)))
        out.append(gates.CU1(int(a), int(c), 0.41))
        out.append(gates.Z(int(b)).controlled_by(int(a)))
        out.append(gates.CCZ(int(a), int(b), int(c)))
        out.append(gates.T(int(c)))
        out.append(gates.Y(int(a)).controlled_by(int(c)))
    return [g for g in out if len(set(g.qubits)) == len(g.qubits)]


@pytest.mark.parametrize("n,low,max_high", [(6, 2, 2), (9, 3, 4), (10, 5, 5), (8, 1, 6)])
def test_tile_plan_matches_gate_by_gate(n, low, max_high):
    rng = np.random.default_rng(n * 7 + low)
    circuit = _random_circuit(n, 4, rng)
    psi = gm.random_state(n, rng)
    want = psi.copy()
    for g in circuit:
        want = port.PortBackend(2).apply_gate(g, want, n)
    got, passes = _run_planned(circuit, psi, n, low, max_high)
    assert len(passes) < len(circuit)
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_qft_phase_ladders_merge():
    n = 12
    c = []
    for i in range(n):
        c.append(gates.H(i))
        for j in range(i + 1, n):
            c.append(gates.CU1(j, i, np.pi / 2 ** (j - i)))
    for i in range(n // 2):
        c.append(gates.SWAP(i, n - i - 1))
    rng = np.random.default_rng(3)
    psi = gm.random_state(n, rng)
    want = psi.copy()
    for g in c:
        want = port.PortBackend(2).apply_gate(g, want, n)
    got, passes = _run_planned(c, psi, n, 4, 5)
    np.testing.assert_allclose(got, want, atol=1e-12)
    nphase = sum(isinstance(op, tiles.PhaseOp) for p in passes if isinstance(p, tiles.Pass) for op in p.ops)
    assert nphase <= n  # one merged ladder per H
    # the final swaps (bit b <-> bit n-1-b) are one bit-swap pass
    swaps = [p for p in passes if isinstance(p, tiles.SwapPass)]
    assert len(swaps) == 1 and sorted(map(sorted, swaps[0].pairs)) == [[b, n - 1 - b] for b in range(n // 2)]
    # closed form: QFT matrix on the bit-reversed... compare with the DFT
    dft = np.fft.ifft(psi) * np.sqrt(2**n)
    np.testing.assert_allclose(got, dft, atol=1e-12)


def test_tile_plan_density_matrix():
    n = 4
    rng = np.random.default_rng(5)
    circuit = _random_circuit(n, 2, rng)
    rho = gm.random_density_matrix(n, rng)
    want = rho.copy()
    for g in circuit:
        want = port.PortBackend(2).apply_gate(g, want, n)
    got, _ = _run_planned(circuit, rho.reshape(-1), 2 * n, 3, 4, density_nq=n)
    np.testing.assert_allclose(got.reshape(rho.shape), want, atol=1e-12)


def test_fanout_expands_to_cnots():
    n = 5
    rng = np.random.default_rng(6)
    psi = gm.random_state(n, rng)
    g = gates.FanOut(1, 0, 3, 4)
    want = port.PortBackend(2).apply_gate(g, psi.copy(), n)
    got, _ = _run_planned([g], psi, n, 2, 3)
    np.testing.assert_allclose(got, want, atol=1e-14)


def test_phase_merge_rules():
    a = tiles.PhaseOp({3}, 0.0, {5: 0.2})
    b = tiles.PhaseOp({4}, 0.0, {5: 0.3})  # shares bit 5 -> merges as control 5
    m = a.merged(b)
    assert m is not None and m.controls == {5} and m.thetas == {3: 0.2, 4: 0.3}
    c = tiles.PhaseOp(set(), -0.1, {1: 0.2})
    d = tiles.PhaseOp(set(), 0.05, {1: 0.1, 2: 0.4})
    m = c.merged(d)
    assert m.controls == frozenset() and abs(m.c0 + 0.05) < 1e-15 and m.thetas == {1: 0.30000000000000004, 2: 0.4}
    assert tiles.PhaseOp({1, 2}, 0.0, {3: 0.1}).merged(tiles.PhaseOp({1}, 0.0, {3: 0.1})) is None


def test_diagonal_gates_become_phase_ops():
    for g in (gates.Z(0), gates.RZ(1, 0.3), gates.CU1(0, 2, 0.5), gates.CZ(1, 2), gates.T(0), gates.CCZ(0, 1, 2)):
        assert isinstance(tiles.gate_to_op(port.custom_matrix, g, 3), tiles.PhaseOp), g.name
    for g in (gates.H(0), gates.X(1), gates.CNOT(0, 1), gates.SWAP(0, 2).controlled_by(1),
              gates.fSim(0, 1, 0.1, 0.2)):
        assert isinstance(tiles.gate_to_op(port.custom_matrix, g, 3), tiles.DenseOp), g.name
    op = tiles.gate_to_op(port.custom_matrix, gates.SWAP(0, 2), 3)
    assert isinstance(op, tiles.SwapOp) and op.targets == [2, 0]
    psi = gm.random_state(3, np.random.default_rng(1))
    got, _ = _run_planned([gates.RZ(1, 0.3)], psi, 3, 1, 2)
    np.testing.assert_allclose(got, er.apply_gate(psi, gm.RZ(0.3), [1]), atol=1e-15)


def test_fixed_matrix_shortcuts_match_the_matrix_path():
    """X / CNOT / TOFFOLI, Z / CZ / CCZ and SWAP skip the matrix encoding in
    gate_to_op: their ops equal the ones built from the dense target block
    (pinned by running both through the emulator against the port)."""
    rng = np.random.default_rng(5)
    n = 5
    cases_ = [gates.X(1), gates.CNOT(3, 0), gates.TOFFOLI(0, 4, 2), gates.Z(2), gates.CZ(4, 1),
              gates.CCZ(0, 1, 3), gates.SWAP(1, 3), gates.SWAP(0, 4).controlled_by(2), gates.X(3).controlled_by(0, 1)]
    for g in cases_:
        op = tiles.gate_to_op(port.custom_matrix, g, n)
        u = tiles.target_block(port.custom_matrix, g)
        if isinstance(op, tiles.PhaseOp):
            np.testing.assert_allclose(np.diag(u), [1, np.exp(1j * sum(op.thetas.values()))], atol=1e-15)
            assert op.c0 == 0.0 and list(op.thetas) == [n - 1 - g.target_qubits[0]]
        elif isinstance(op, tiles.DenseOp):
            np.testing.assert_array_equal(op.matrix, u)
        assert sorted(op.controls) == sorted(n - 1 - q for q in g.control_qubits)
        psi = gm.random_state(n, rng)
        got, _ = _run_planned([g], psi, n, 2, 3)
        want = port.PortBackend(1).apply_gate(g, psi.copy(), n)
        np.testing.assert_allclose(got, want, atol=1e-14)


def test_plan_blocks_respects_dependencies_and_width():
    import itertools

    from qibojit_amd.fusion import plan_blocks

    rng = np.random.default_rng(5)
    n = 9
    seq = [tuple(rng.choice(n, size=int(rng.integers(1, 3)), replace=False).tolist()) for _ in range(200)]
    blocks = plan_blocks(seq, 5)
    order = [i for _, idx in blocks for i in idx]
    assert sorted(order) == list(range(len(seq)))
    pos = {i: k for k, i in enumerate(order)}
    # every pair of gates sharing a qubit keeps its relative order
    for i, j in itertools.combinations(range(len(seq)), 2):
        if set(seq[i]) & set(seq[j]):
            assert pos[i] < pos[j]
    for qubits, idx in blocks:
        assert len(qubits) <= 5 and set().union(*(set(seq[i]) for i in idx)) == set(qubits)


def test_plan_blocks_fuses_a_brickwork():
    from qibojit_amd.fusion import plan_blocks

    seq = [(q, q + 1) for d in range(40) for q in range(d % 2, 31, 2)]
    blocks = plan_blocks(seq, 5)
    assert len(blocks) <= len(seq) / 4
    assert max(len(q) for q, _ in blocks) == 5


def test_swap_runs_become_bit_swap_passes():
    """In-tile swaps relabel bits; a run of swaps on distinct qubits that
    does not fit the pass is one bit-swap pass (emulated on the CPU)."""
    n = 11
    rng = np.random.default_rng(21)
    c = []
    for q in range(n):
        c.append(gates.U3(q, *rng.uniform(0, 2 * np.pi, 3)))
    c += [gates.SWAP(0, 10), gates.SWAP(1, 9), gates.SWAP(2, 3)]          # run of 3, spans the state
    c += [gates.CU1(4, 0, 0.3), gates.SWAP(5, 6), gates.H(6)]            # lone swap: in a tile pass
    c += [gates.SWAP(7, 8), gates.SWAP(8, 9), gates.RX(8, 0.4)]           # overlapping: two runs of 1
    c += [gates.SWAP(0, 1).controlled_by(4), gates.CZ(2, 9)]              # controlled swap stays dense
    psi = gm.random_state(n, rng)
    want = psi.copy()
    for g in c:
        want = port.PortBackend(2).apply_gate(g, want, n)
    got, passes = _run_planned(c, psi, n, 3, 4)
    np.testing.assert_allclose(got, want, atol=1e-12)
    assert any(isinstance(p, tiles.SwapPass) and len(p.pairs) == 3 for p in passes)
    in_tile = [op for p in passes if isinstance(p, tiles.Pass) for op in p.ops if isinstance(op, tiles.SwapOp)]
    assert len(in_tile) >= 2


def test_bit_swaps_emulator_closed_form():
    """apply_bit_swaps == numpy axis permutation (qubit q = axis q, bit n-1-q)."""
    n = 7
    psi = gm.random_state(n, np.random.default_rng(8))
    pairs = [(0, 6), (2, 4)]
    perm = list(range(n))
    for a, b in pairs:
        qa, qb = n - 1 - a, n - 1 - b
        perm[qa], perm[qb] = perm[qb], perm[qa]
    want = np.transpose(psi.reshape((2,) * n), perm).reshape(-1)
    np.testing.assert_allclose(apply_bit_swaps(psi, n, pairs), want, atol=0)
