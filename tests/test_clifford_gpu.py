"""Stabilizer-tableau kernels on the GPU against the oracle, bit for bit.

Gate rules, layout conversion, gate lists in one launch (LDS and global
variants), rowsum, the random and determined outcomes and the measurement
loop are compared with oracle/clifford_ref.py on the same matrices; at sizes
the oracle cannot hold, a gate list followed by its inverse must return the
tableau unchanged, and a repeated measurement must repeat its outcome.
"""

import os

import numpy as np
import pytest
import torch

from oracle import clifford_ref as cr
from qibojit_amd.backends import clifford_operations_hip as ch
from tests.test_clifford import random_circuit, REFERENCE_DETERMINED_CASE

pytestmark = pytest.mark.gpu

NAMES1 = list(cr.ONE_QUBIT)
NAMES2 = list(cr.TWO_QUBIT)
INVERSE = {"H": ["H"], "S": ["SDG"], "SDG": ["S"], "SX": ["SXDG"], "SXDG": ["SX"], "X": ["X"], "Y": ["Y"],
           "Z": ["Z"], "RY_pi": ["RY_3pi_2"], "RY_3pi_2": ["RY_pi"], "CNOT": ["CNOT"], "CZ": ["CZ"],
           "SWAP": ["SWAP"], "CY": ["CY"], "iSWAP": ["iSWAP"] * 3}


def _dense(rng, n):
    return rng.integers(0, 2, size=(2 * n + 1, 2 * n + 1)).astype(np.uint8)


def _to_host(tab, n):
    return ch._clifford_post_execution_reshape(tab, n).cpu().numpy()


def _random_gates(rng, n, ngates):
    gates = []
    for _ in range(ngates):
        if n >= 2 and rng.random() < 0.5:
            gates.append((NAMES2[rng.integers(len(NAMES2))], *[int(v) for v in rng.choice(n, 2, replace=False)]))
        else:
            gates.append((NAMES1[rng.integers(len(NAMES1))], int(rng.integers(n))))
    return gates


def _oracle_run(sm, gates, n):
    for name, *qs in gates:
        cr.apply(sm, name, tuple(qs), n)
    return sm


@pytest.mark.parametrize("n", [1, 15, 16, 47, 48, 200])
def test_pack_unpack_round_trip(n):
    rng = np.random.default_rng(n)
    sm = _dense(rng, n)
    tab = ch._clifford_pre_execution_reshape(sm)
    assert tab.shape == (2 * n + 1, ch.column_words(n))
    np.testing.assert_array_equal(_to_host(tab, n), sm)
    # padding rows and words stay zero
    words = tab.cpu().numpy().view(np.uint32)
    nrow = 2 * n + 1
    for c in range(0, nrow, max(1, nrow // 7)):
        bits = np.unpackbits(words[c].view(np.uint8), bitorder="little")
        assert not bits[nrow:].any()


@pytest.mark.parametrize("name", NAMES1 + NAMES2)
@pytest.mark.parametrize("n", [2, 40, 130])
def test_single_gate(name, n):
    rng = np.random.default_rng(abs(hash((name, n))) % 2**32)
    for _ in range(3):
        sm = _dense(rng, n)
        qs = tuple(int(v) for v in rng.choice(n, size=2 if name in NAMES2 else 1, replace=False))
        tab = ch._clifford_pre_execution_reshape(sm)
        out = getattr(ch, name)(tab, *qs, n)
        assert out is tab
        np.testing.assert_array_equal(_to_host(tab, n), cr.apply(sm.copy(), name, qs, n), err_msg=f"{name} {qs}")


@pytest.mark.parametrize("kernel", ["lds1", "lds4", "lds8", "lds16", "global1", "global4", "global8", "global16"])
@pytest.mark.parametrize("n,ngates", [(1, 50), (3, 400), (17, 600), (100, 1500), (1000, 600)])
def test_circuit_matches_gate_by_gate(n, ngates, kernel, monkeypatch):
    """Every gate-list variant: columns in an LDS slice or in L2, with 1, 4,
    8 or 16 waves per workgroup sharing each group's slots."""
    if kernel.startswith("global"):
        monkeypatch.setenv("QJ_CLIFFORD_KERNEL", "global")
    monkeypatch.setenv("QJ_CLIFFORD_WAVES", kernel.lstrip("ldsglobal"))
    rng = np.random.default_rng(1000 * n + ngates)
    sm = _dense(rng, n)
    gates = _random_gates(rng, n, ngates)
    tab = ch._clifford_pre_execution_reshape(sm)
    ch.apply_circuit(tab, gates, n)
    np.testing.assert_array_equal(_to_host(tab, n), _oracle_run(sm.copy(), gates, n))


@pytest.mark.parametrize("waves", ["1", "4", "8", "16"])
def test_compiled_schedule_applied_twice(waves, monkeypatch):
    monkeypatch.setenv("QJ_CLIFFORD_WAVES", waves)
    rng = np.random.default_rng(44)
    n = 60
    gates = _random_gates(rng, n, 2000)
    sm = _dense(rng, n)
    sched = ch.compile_circuit(gates, n)
    assert sched.ngates == 2000 and sched.ngroups >= 1
    tab = ch._clifford_pre_execution_reshape(sm)
    ch.apply_schedule(tab, sched, n)
    ch.apply_schedule(tab, sched, n)
    want = _oracle_run(_oracle_run(sm.copy(), gates, n), gates, n)
    np.testing.assert_array_equal(_to_host(tab, n), want)
    with pytest.raises(ValueError, match="compiled for"):
        ch.apply_schedule(ch.zero_state(n + 1), sched, n + 1)


@pytest.mark.parametrize("waves", ["1", "4", "8", "16"])
def test_circuit_of_dependent_chain(waves, monkeypatch):
    """Gates that all share a qubit: one gate per group, order kept."""
    monkeypatch.setenv("QJ_CLIFFORD_WAVES", waves)
    rng = np.random.default_rng(3)
    n = 6
    gates = []
    for _ in range(300):
        if rng.random() < 0.5:
            gates.append((NAMES2[rng.integers(len(NAMES2))], 0, int(rng.integers(1, n))))
        else:
            gates.append((NAMES1[rng.integers(len(NAMES1))], 0))
    sm = _dense(rng, n)
    tab = ch._clifford_pre_execution_reshape(sm)
    ch.apply_circuit(tab, gates, n)
    np.testing.assert_array_equal(_to_host(tab, n), _oracle_run(sm.copy(), gates, n))


def _random_device_tableau(n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    ws = ch.column_words(n)
    tab = torch.randint(-(2**31), 2**31 - 1, (2 * n + 1, ws), dtype=torch.int32, device="cuda", generator=g)
    nrow = 2 * n + 1
    full, rem = divmod(nrow, 32)
    if rem:
        tab[:, full] &= (1 << rem) - 1
        tab[:, full + 1:] = 0
    else:
        tab[:, full:] = 0
    return tab


@pytest.mark.parametrize("n", [20500])
def test_circuit_then_inverse_is_identity_at_global_size(n):
    """n = 20500: a column does not fit the LDS slice (global variant)."""
    rng = np.random.default_rng(n)
    gates = _random_gates(rng, n, 3000)
    inverse = [(g, *qs) for name, *qs in reversed(gates) for g in INVERSE[name]]
    tab = _random_device_tableau(n, 7)
    before = tab.clone()
    ch.apply_circuit(tab, gates, n)
    assert not torch.equal(tab, before)
    ch.apply_circuit(tab, inverse, n)
    assert torch.equal(tab, before)


def test_rowsum_parity():
    rng = np.random.default_rng(21)
    for n in (3, 40, 150):
        for determined in (False, True):
            sm = _dense(rng, n)
            rows = rng.permutation(2 * n + 1)
            m = int(rng.integers(1, n + 1))
            if determined:
                h = [int(rows[0])] * m
                i = [int(v) for v in rng.choice(rows[1:], size=m)]
            else:
                h = [int(v) for v in rows[:m]]
                i = [int(v) for v in rng.choice(rows[m:], size=m)]
            tab = ch._clifford_pre_execution_reshape(sm)
            ch._rowsum(tab, h, i, n, determined)
            np.testing.assert_array_equal(_to_host(tab, n), cr.rowsum(sm.copy(), h, i, n, determined),
                                          err_msg=f"n={n} determined={determined}")


@pytest.mark.parametrize("n", [2, 31, 100, 700])
def test_outcomes_on_random_matrices(n):
    rng = np.random.default_rng(n + 5)
    for trial in range(4):
        sm = _dense(rng, n)
        q = int(rng.integers(n))
        tab = ch._clifford_pre_execution_reshape(sm)
        p = ch.first_stabilizer(tab, q, n)
        assert p == cr.first_stabilizer(sm, q, n)
        if p >= 0:
            np.random.seed(100 * n + trial)
            _, outcome = ch._random_outcome(tab, [p], q, n)
            want, _ = cr.random_outcome(sm.copy(), p, q, n, int(outcome))
        else:
            _, outcome = ch._determined_outcome(tab, q, n)
            want, wout = cr.determined_outcome(sm.copy(), q, n)
            assert int(outcome) == wout
        np.testing.assert_array_equal(_to_host(tab, n), want, err_msg=f"trial {trial} q={q} p={p}")
        # the determined rule on the same matrix, whatever p is
        tab = ch._clifford_pre_execution_reshape(sm)
        _, outcome = ch._determined_outcome(tab, q, n)
        want, wout = cr.determined_outcome(sm.copy(), q, n)
        assert int(outcome) == wout
        np.testing.assert_array_equal(_to_host(tab, n), want)


def _oracle_measure(sm, qubits, n, draws):
    sample = []
    for q in qubits:
        p = cr.first_stabilizer(sm, q, n)
        if p >= 0:
            sm, out = cr.random_outcome(sm, p, q, n, int(next(draws)))
        else:
            sm, out = cr.determined_outcome(sm, q, n)
        sample.append(out)
    return sm, sample


@pytest.mark.parametrize("n", [3, 12, 64])
def test_measurement_loop(n):
    rng = np.random.default_rng(n)
    gates = random_circuit(rng, n, 8 * n)
    sm = _oracle_run(cr.zero_tableau(n), gates, n)
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, gates, n)
    np.testing.assert_array_equal(_to_host(tab, n), sm)
    qubits = list(rng.permutation(n))
    np.random.seed(n)
    sample = ch.M(tab, qubits, n, collapse=True)
    np.random.seed(n)
    draws = iter(lambda: np.random.randint(2, size=None, dtype=np.uint), None)
    want_sm, want = _oracle_measure(sm.copy(), qubits, n, draws)
    assert sample == want
    np.testing.assert_array_equal(_to_host(tab, n), want_sm)


def _per_qubit_loop(tab, qubits, n, exact=False):
    sample, nrandom = [], 0
    for q in qubits:
        p = ch.first_stabilizer(tab, q, n)
        if p >= 0:
            _, out = ch._random_outcome(tab, [p], q, n)
            nrandom += 1
        else:
            _, out = ch._determined_outcome(tab, q, n, exact)
        sample.append(int(out))
    return sample, nrandom


@pytest.mark.parametrize("path", ["auto", "kernels", "small"])
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("n,nmeas,ngates", [(1, 1, 3), (7, 14, 40), (33, 40, 400), (256, 100, 3000),
                                            (300, 120, 3000), (1000, 200, 20000)])
def test_device_measurement_loop_equals_per_qubit_calls(n, nmeas, ngates, exact, path, monkeypatch):
    """M's one-call device loop against the per-qubit entry points: same
    outcomes, same collapsed tableau, and the generator left where the
    per-qubit loop's scalar draws leave it (qubits repeat when nmeas > n).
    Up to 256 qubits the loop runs in one workgroup unless the per-step
    kernels are forced; the one-workgroup loop can be forced up to 1024."""
    if path == "kernels":
        if n > 256:
            pytest.skip("the per-step kernels are the default above 256 qubits")
        monkeypatch.setenv("QJ_CLIFFORD_MEASURE", "kernels")
    if path == "small":
        if n <= 256:
            pytest.skip("the one-workgroup loop is the default up to 256 qubits")
        monkeypatch.setenv("QJ_CLIFFORD_MEASURE", "small")
    rng = np.random.default_rng(n + 3)
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, _random_gates(rng, n, ngates), n)
    qubits = [int(v) for v in rng.integers(0, n, size=nmeas)]
    ref = tab.clone()
    np.random.seed(n)
    got = ch.M(tab, qubits, n, collapse=True, exact=exact)
    after = np.random.randint(2**31)
    np.random.seed(n)
    want, nrandom = _per_qubit_loop(ref, qubits, n, exact)
    assert got == want
    assert after == np.random.randint(2**31)
    assert torch.equal(tab, ref)
    if n > 1:
        assert 0 < nrandom < nmeas  # both branches taken


def test_device_measurement_loop_with_no_qubits():
    tab = ch.zero_state(4)
    before = tab.clone()
    np.random.seed(1)
    assert ch.M(tab, [], 4, collapse=True) == []
    assert np.random.randint(2**31) == np.random.RandomState(1).randint(2**31)
    assert torch.equal(tab, before)


def test_non_collapsing_measurement_keeps_the_state():
    n = 10
    rng = np.random.default_rng(2)
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, random_circuit(rng, n, 60), n)
    before = tab.clone()
    ch.M(tab, list(range(n)), n)
    assert torch.equal(tab, before)


@pytest.mark.parametrize("n,nmeas", [(1, 1), (5, 5), (24, 20), (90, 40)])
def test_sample_shots_equals_shot_by_shot(n, nmeas):
    """Affine-map sampling == nshots measurement loops on copies, draw for draw."""
    rng = np.random.default_rng(n)
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, random_circuit(rng, n, 10 * n), n)
    qubits = [int(v) for v in rng.permutation(n)[:nmeas]]
    before = tab.clone()
    np.random.seed(7)
    got = ch.sample_shots(tab, qubits, n, 25).cpu().numpy()
    _, keys, pos, *_ = np.random.get_state()
    assert torch.equal(tab, before)
    np.random.seed(7)
    want = np.array([ch.M(tab, qubits, n) for _ in range(25)], dtype=np.uint8)
    np.testing.assert_array_equal(got, want)
    _, keys2, pos2, *_ = np.random.get_state()
    assert np.array_equal(keys, keys2) and pos == pos2


@pytest.mark.parametrize("n,nmeas", [(6, 6), (30, 18)])
def test_measurement_forms_against_forced_outcomes(n, nmeas):
    """The symbolic pass's forms equal concrete measurement runs with the
    random outcomes forced: all 0, and each one alone set to 1."""
    rng = np.random.default_rng(100 + n)
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, random_circuit(rng, n, 12 * n), n)
    qubits = [int(v) for v in rng.permutation(n)[:nmeas]]
    forms, k = ch.measurement_forms(tab, qubits, n)
    assert forms.shape == (nmeas, k + 1) and k >= 2
    s0, at = ch.measure_with_outcomes(tab.clone(), qubits, n, set())
    assert len(at) == k
    np.testing.assert_array_equal(forms[:, 0], s0)
    for j in range(k):
        sj, _ = ch.measure_with_outcomes(tab.clone(), qubits, n, {j})
        np.testing.assert_array_equal(forms[:, 0] ^ forms[:, j + 1], sj, err_msg=f"variable {j + 1}")


def test_sample_shots_of_a_ghz_state():
    """GHZ on 40 qubits: every shot is all zeros or all ones, about half each."""
    n = 40
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, [("H", 0)] + [("CNOT", 0, q) for q in range(1, n)], n)
    np.random.seed(3)
    shots = ch.sample_shots(tab, list(range(n)), n, 4000).cpu().numpy()
    assert np.all((shots.sum(axis=1) == 0) | (shots.sum(axis=1) == n))
    assert 1800 < int(shots[:, 0].sum()) < 2200


def test_reference_determined_rule_case():
    n, gates, q = REFERENCE_DETERMINED_CASE
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, gates, n)
    assert ch.first_stabilizer(tab, q, n) == -1
    _, outcome = ch._determined_outcome(tab, q, n)
    assert int(outcome) == 1  # the reference's rule (the state's certain outcome is 0)


@pytest.mark.parametrize("n", [1, 2, 31, 33, 100, 700, 2100])
def test_exact_determined_outcome_on_random_matrices(n):
    """The exact rule bit for bit against the oracle on random 0/1 matrices
    (any selection, so long products and both outcomes occur); n = 2100
    spans more than one 64-word pass of a column."""
    rng = np.random.default_rng(n + 50)
    outcomes = set()
    for trial in range(6 if n < 2100 else 3):
        sm = _dense(rng, n)
        q = int(rng.integers(n))
        tab = ch._clifford_pre_execution_reshape(sm)
        _, outcome = ch._determined_outcome(tab, q, n, exact=True)
        want, wout = cr.determined_outcome_exact(sm.copy(), q, n)
        assert int(outcome) == wout, f"trial {trial} q={q}"
        np.testing.assert_array_equal(_to_host(tab, n), want, err_msg=f"trial {trial} q={q}")
        outcomes.add(wout)
    if n >= 31:
        assert outcomes == {0, 1}


def test_exact_rule_on_the_reference_case():
    n, gates, q = REFERENCE_DETERMINED_CASE
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, gates, n)
    _, outcome = ch._determined_outcome(tab, q, n, exact=True)
    assert int(outcome) == 0  # the state's certain outcome
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, gates, n)
    assert ch.M(tab, [q], n, exact=True) == [0]
    assert ch.sample_shots(tab, [q], n, 8, exact=True).cpu().numpy().tolist() == [[0]] * 8


@pytest.mark.parametrize("n,nmeas", [(12, 12), (90, 40)])
def test_exact_measurement_loop_and_sampling(n, nmeas):
    """M(exact=True) against the oracle's loop with the exact rule, and
    sample_shots(exact=True) against M(exact=True) shot by shot."""
    rng = np.random.default_rng(n + 1)
    gates = random_circuit(rng, n, 10 * n)
    sm = _oracle_run(cr.zero_tableau(n), gates, n)
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, gates, n)
    qubits = [int(v) for v in rng.permutation(n)[:nmeas]]
    before = tab.clone()
    np.random.seed(n)
    got = ch.sample_shots(tab, qubits, n, 20, exact=True).cpu().numpy()
    assert torch.equal(tab, before)
    np.random.seed(n)
    want = np.array([ch.M(tab, qubits, n, exact=True) for _ in range(20)], dtype=np.uint8)
    np.testing.assert_array_equal(got, want)
    np.random.seed(n)
    sample = ch.M(tab, qubits, n, collapse=True, exact=True)
    np.random.seed(n)
    osm, osample = sm.copy(), []
    for q in qubits:
        p = cr.first_stabilizer(osm, q, n)
        if p >= 0:
            osm, out = cr.random_outcome(osm, p, q, n, int(np.random.randint(2, size=None, dtype=np.uint)))
        else:
            osm, out = cr.determined_outcome_exact(osm, q, n)
        osample.append(out)
    assert sample == osample
    np.testing.assert_array_equal(_to_host(tab, n), osm)


def test_repeated_measurement_at_size():
    """n = 4096: measuring every qubit twice gives the same outcomes."""
    n = 4096
    rng = np.random.default_rng(9)
    tab = ch.zero_state(n)
    ch.apply_circuit(tab, _random_gates(rng, n, 40000), n)
    qubits = [int(v) for v in rng.choice(n, size=24, replace=False)]
    np.random.seed(4)
    first = ch.M(tab, qubits, n, collapse=True)
    again = ch.M(tab, qubits, n, collapse=True)
    assert first == again


@pytest.mark.parametrize("rows,cols", [(1, 1), (5, 8), (7, 13), (33, 64), (3, 200)])
def test_packbits_like_numpy(rows, cols):
    rng = np.random.default_rng(rows * cols)
    a = rng.integers(0, 2, size=(rows, cols)).astype(np.uint8)
    for axis in (0, 1):
        got = ch._packbits(a, axis).cpu().numpy()
        np.testing.assert_array_equal(got, np.packbits(a, axis=axis))
        count = a.shape[axis]
        back = ch._unpackbits(got, axis, count).cpu().numpy()
        np.testing.assert_array_equal(back, np.unpackbits(np.packbits(a, axis=axis), axis=axis, count=count))
        np.testing.assert_array_equal(back, a)


def test_pack_for_measurements_round_trip():
    rng = np.random.default_rng(8)
    n = 11
    sm = _dense(rng, n)
    packed = ch._pack_for_measurements(sm, n).cpu().numpy()
    want = np.hstack((np.packbits(sm[:, :n], axis=1), np.packbits(sm[:, n:-1], axis=1), sm[:, -1:]))
    np.testing.assert_array_equal(packed, want)
    np.testing.assert_array_equal(ch._unpack_for_measurements(packed, n).cpu().numpy(), sm)


def test_bad_tableau_is_rejected():
    tab = ch.zero_state(4)
    with pytest.raises(ValueError, match="does not match"):
        ch.H(tab, 0, 5)
    with pytest.raises(TypeError):
        ch.H(tab.cpu(), 0, 4)
    with pytest.raises(ValueError, match="out of range"):
        ch.CNOT(tab, 0, 4, 4)
    assert os.environ.get("QJ_CLIFFORD_KERNEL") is None
