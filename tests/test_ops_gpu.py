"""State preparation, collapse, probabilities, Metropolis sampling and FP64
reductions on the GPU, against the CPU oracle (collapse and sampling
indices bit-exact)."""

import json
import os

import numpy as np
import pytest

from oracle import einsum_ref as er
from oracle import gate_matrices as gm
from oracle import port
from tests import cases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("is_matrix", [False, True])
def test_zero_state(backend, dtype, is_matrix):
    """tests/test_ops.py:12-22."""
    backend.set_dtype(dtype)
    if is_matrix:
        got = backend.zero_state(4, density_matrix=True)
        want = np.reshape(np.array([1] + [0] * 255, dtype=dtype), (16, 16))
    else:
        got = backend.zero_state(4)
        want = np.array([1] + [0] * 15, dtype=dtype)
    got = backend.to_numpy(got)
    assert got.dtype == np.dtype(dtype) and got.shape == want.shape
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n,targets,results", cases.COLLAPSE)
@pytest.mark.parametrize("normalize", [True, False])
def test_collapse_state(backend, n, targets, results, normalize, dtype, rng):
    """tests/test_ops.py:52-83; kept index set bit-exact vs the port."""
    backend.set_dtype(dtype)
    psi = gm.random_state(n, rng, dtype)
    b2d = 2 ** np.arange(len(results) - 1, -1, -1)
    result = np.array([np.array(results).dot(b2d)])
    got = backend.to_numpy(backend.collapse_state(backend.cast(psi, dtype=dtype, copy=True), sorted(targets), result, n, normalize))
    want = port.PortBackend(4).collapse_state(psi.copy(), sorted(targets), int(result[0]), n, normalize)
    np.testing.assert_array_equal(got == 0, want == 0)
    np.testing.assert_allclose(got, want, atol=1e-14 if dtype == "complex128" else 1e-7, rtol=0)
    np.testing.assert_allclose(got, er.collapse(psi, targets, results, normalize), atol=1e-14 if dtype == "complex128" else 1e-7)


def test_collapse_unsorted_qubit_order(backend, rng):
    """Unsorted measured qubits: the first listed qubit is the most
    significant bit of the outcome.  (The reference's collapse_index shifts
    earlier-inserted bits past the vector for such input, so the CPU port is
    only ever called with sorted qubits.)"""
    n = 7
    psi = gm.random_state(n, rng)
    for qubits, shot in (([5, 1, 3], 5), ([0, 6], 2), ([4, 2, 6, 0], 9)):
        got = backend.to_numpy(backend.collapse_state(backend.cast(psi, copy=True), qubits, shot, n, False))
        results = [(shot >> (len(qubits) - 1 - i)) & 1 for i in range(len(qubits))]
        np.testing.assert_array_equal(got, er.collapse(psi, qubits, results, normalize=False))


@pytest.mark.parametrize("qubits", [[0], [3], [1, 4], [4, 1], [5, 0, 2], [0, 1, 2, 3, 4], [2, 0, 5, 1, 4, 3], list(range(9))])
def test_calculate_probabilities(backend, qubits, dtype, rng):
    backend.set_dtype(dtype)
    n = 9
    psi = gm.random_state(n, rng, dtype)
    got = backend.to_numpy(backend.calculate_probabilities(backend.cast(psi, dtype=dtype), qubits, n))
    assert got.dtype == (np.float32 if dtype == "complex64" else np.float64)
    np.testing.assert_allclose(got, er.probabilities(psi, qubits), atol=1e-6 if dtype == "complex64" else 1e-14)


def test_probabilities_deterministic(backend, rng):
    psi = backend.cast(gm.random_state(22, rng))
    a = backend.to_numpy(backend.calculate_probabilities(psi, [0, 7, 21], 22))
    b = backend.to_numpy(backend.calculate_probabilities(psi, [0, 7, 21], 22))
    np.testing.assert_array_equal(a, b)
    assert abs(a.sum() - 1) < 1e-12


@pytest.mark.parametrize("realtype", ["float32", "float64"])
@pytest.mark.parametrize("inttype", ["int32", "int64"])
def test_measure_frequencies_golden(backend, realtype, inttype):
    """tests/test_ops.py:236-256 golden vector."""
    with open(os.path.join(GOLDEN, "metropolis_uniform16.json"), encoding="utf-8") as f:
        gold = json.load(f)
    probs = np.ones(16, dtype=realtype) / 16
    freqs = np.zeros(16, dtype=inttype)
    out = backend.measure_frequencies_op(freqs, probs, nshots=1000, nqubits=4, seed=1234, nthreads=4)
    assert out.sum() == 1000
    np.testing.assert_array_equal(out, np.array(gold["frequencies"], dtype=inttype))


@pytest.mark.parametrize(
    "nstates,nthreads,nshots",
    [(2, 1, 5000), (8, 4, 100000), (8, 16, 1000000), (16, 64, 777777), (64, 3, 200003), (5, 4, 30001), (128, 8, 50000), (1, 2, 100)],
)
@pytest.mark.parametrize("realtype", ["float32", "float64"])
def test_measure_frequencies_bitexact_vs_port(backend, nstates, nthreads, nshots, realtype):
    rng = np.random.default_rng(nstates + 1000 * nthreads)
    probs = rng.random(nstates).astype(realtype)
    if nstates > 2:
        probs[rng.integers(nstates)] = 0
    probs = (probs / probs.sum()).astype(realtype)
    seed = int(rng.integers(0, 10**8))
    got = backend.measure_frequencies_op(np.zeros(nstates, dtype=np.int64), probs, nshots, 0, seed, nthreads)
    want = port.measure_frequencies(np.zeros(nstates, dtype=np.int64), probs, nshots, 0, seed, nthreads)
    np.testing.assert_array_equal(got, want)


def test_sample_frequencies_sparse_probabilities(backend):
    """tests/test_ops.py:259-275 (both sides of the Metropolis threshold)."""
    import itertools

    nonzero = []
    for r in (1, 2, 3, 4):
        nonzero.extend(itertools.combinations(range(8), r))
    for i, nz in enumerate(nonzero):
        probs = np.zeros(8)
        probs[list(nz)] = 1
        probs = probs / probs.sum()
        nshots = 200000 if i % 2 else 1000
        freqs = backend.sample_frequencies(probs, nshots)
        assert sum(freqs.values()) == nshots
        for k, v in freqs.items():
            assert v != 0 and k in nz


def test_sample_frequencies_uses_seeded_chains(backend):
    backend.set_seed(42)
    probs = np.array([0.1, 0.2, 0.3, 0.4])
    a = backend.sample_frequencies(probs, 300000)
    backend.set_seed(42)
    seed = np.random.RandomState(42).randint(0, int(1e8), size=1)[0]
    want = port.measure_frequencies(np.zeros(4, dtype=np.int64), probs, 300000, 2, seed, backend.nthreads)
    assert dict(a) == {i: int(f) for i, f in enumerate(want) if f > 0}


def test_reductions(backend, dtype, rng):
    backend.set_dtype(dtype)
    n = 12
    a = gm.random_state(n, rng, dtype)
    b = gm.random_state(n, rng, dtype)
    da, db = backend.cast(a, dtype=dtype), backend.cast(b, dtype=dtype)
    tol = 1e-13 if dtype == "complex128" else 1e-6
    assert abs(backend.norm_squared(da) - np.vdot(a.astype(complex), a.astype(complex)).real) < tol
    assert abs(backend.vdot(da, db) - np.vdot(a.astype(complex), b.astype(complex))) < tol
    for g in (1.0, 0.35):
        assert abs(backend.expectation_tfim(da, n, g) - er.tfim_energy(a, g)) < tol * n


@pytest.mark.parametrize("size,offset", [(1001, 0), (1000, 1), (4097, 1), (3, 0)])
def test_reductions_odd_sizes_and_offsets(backend, dtype, rng, size, offset):
    """Norm and overlap over vectors of odd length and from an amplitude
    offset (complex64 then starts off the 16-byte grid the kernel's vector
    loads use): the one-amplitude path and the tail are exact sums too."""
    a = (rng.normal(size=size + offset) + 1j * rng.normal(size=size + offset)).astype(dtype)
    b = (rng.normal(size=size + offset) + 1j * rng.normal(size=size + offset)).astype(dtype)
    da, db = backend.cast(a, dtype=dtype)[offset:], backend.cast(b, dtype=dtype)[offset:]
    a, b = a[offset:].astype(complex), b[offset:].astype(complex)
    # the products are exact in FP64 for both dtypes; only the summation
    # order differs from NumPy's
    want_n, want_v = np.vdot(a, a).real, np.vdot(a, b)
    assert abs(backend.norm_squared(da) - want_n) < 1e-12 * (1 + abs(want_n))
    assert abs(backend.vdot(da, db) - want_v) < 1e-12 * (1 + abs(want_v))


def test_reductions_deterministic(backend, rng):
    a = backend.cast(gm.random_state(24, rng))
    assert backend.norm_squared(a) == backend.norm_squared(a)
    assert backend.expectation_tfim(a, 24) == backend.expectation_tfim(a, 24)


def test_collapse_large_state_idempotent(backend):
    """26-qubit collapse on {0, 7, 25}: normalisation and the collapsed
    marginal are size-independent checks."""
    n = 26
    backend.set_dtype("complex64")
    rng = np.random.default_rng(0)
    psi = backend.cast(gm.random_state(n, rng, np.complex64), dtype="complex64")
    probs = backend.to_numpy(backend.calculate_probabilities(psi, [0, 7, n - 1], n))
    assert abs(probs.sum() - 1) < 1e-5
    out = backend.collapse_state(psi, [0, 7, n - 1], 5, n)
    assert abs(backend.norm_squared(out) - 1) < 1e-5
    p2 = backend.to_numpy(backend.calculate_probabilities(out, [0, 7, n - 1], n))
    np.testing.assert_allclose(p2, np.eye(8)[5], atol=1e-5)


def test_compare_states_mixed_precision(backend, rng):
    from oracle.gate_matrices import random_state

    a = random_state(9, rng)
    b = a + 1e-3 * random_state(9, rng)
    res = backend.compare_states(backend.cast(a.astype(np.complex64), dtype="complex64"), backend.cast(b))
    a64 = a.astype(np.complex64).astype(np.complex128)
    ov = np.vdot(a64, b)
    assert abs(res["overlap"] - ov) < 1e-12
    assert res["norm2_a"] == pytest.approx(np.vdot(a64, a64).real, abs=1e-12)
    assert res["norm2_b"] == pytest.approx(np.vdot(b, b).real, abs=1e-12)
    assert res["max_abs_diff"] == pytest.approx(np.max(np.abs(a64 - b)), rel=1e-12)
    want_inf = 1 - abs(ov) ** 2 / (np.vdot(a64, a64).real * np.vdot(b, b).real)
    assert res["infidelity"] == pytest.approx(want_inf, abs=1e-12)
    same = backend.compare_states(backend.cast(a), backend.cast(a))
    assert same["max_abs_diff"] == 0.0 and abs(same["infidelity"]) < 1e-14


@pytest.mark.parametrize("qubits,shot", [([1], 1), ([0, 3], 2), ([3, 1, 2], 5), ([2, 0], 0)])
def test_density_matrix_probabilities_and_collapse(backend, dtype, qubits, shot):
    rng = np.random.default_rng(17 + shot)
    n = 4
    backend.set_dtype(dtype)
    rho = gm.random_density_matrix(n, rng)
    probs = backend.to_numpy(backend.calculate_probabilities(backend.cast(rho.copy()), qubits, n, density_matrix=True))
    tol = 1e-5 if dtype == "complex64" else 1e-12
    np.testing.assert_allclose(probs, er.probabilities_density_matrix(rho, qubits), atol=tol)
    results = [(shot >> (len(qubits) - 1 - j)) & 1 for j in range(len(qubits))]
    for normalize in (True, False):
        got = backend.to_numpy(backend.collapse_state(backend.cast(rho.copy()), qubits, shot, n, normalize=normalize,
                                                      density_matrix=True))
        want = er.collapse_density_matrix(rho, qubits, results, normalize=normalize)
        np.testing.assert_allclose(got, want, atol=tol)


@pytest.mark.parametrize("nstates", [1, 2, 8, 1000, 2**16, 2**17 + 3])
@pytest.mark.parametrize("realtype", [np.float64, np.float32])
def test_sample_shots_match_numpy_choice(backend, nstates, realtype):
    """sample_shots == RandomState(seed).choice(n, nshots, p) for the same
    seed (the generator the reference's host sampling uses): identical up
    to 2^16 states (sequential FP64 cumsum, as NumPy's); above, the blocked
    scan may round a CDF entry differently, so a sample differs only if a
    uniform falls within that rounding of a boundary (none expected)."""
    rng = np.random.default_rng(nstates)
    p = rng.random(nstates) ** 3
    p[rng.integers(nstates, size=max(1, nstates // 5))] = 0
    if p.sum() == 0:
        p[0] = 1
    p = (p / p.sum()).astype(realtype)
    nshots = 5000
    backend.set_seed(1234)
    got = backend.to_numpy(backend.sample_shots(p, nshots))
    want = np.random.RandomState(1234).choice(nstates, size=nshots, p=p.astype(np.float64))
    assert got.dtype == np.int64 and got.shape == (nshots,)
    np.testing.assert_array_equal(got, want)
    assert np.all(p[got] > 0)
    # the generator advanced exactly as choice() advances it
    rs = np.random.RandomState(1234)
    rs.choice(nstates, size=nshots, p=p.astype(np.float64))
    assert backend._rng.randint(0, 2**31) == rs.randint(0, 2**31)


@pytest.mark.parametrize("nstates", [1000, 2**17 + 3])
@pytest.mark.parametrize("scale", [2.0, 0.5])
@pytest.mark.parametrize("realtype", [np.float64, np.float32])
def test_sample_shots_unnormalized_probabilities(backend, nstates, scale, realtype):
    """Probabilities summing to `scale` sample as choice() on p / sum(p).
    A power-of-two scale changes neither the CDF's rounding nor the quotient
    cdf[i] / cdf[-1], so the samples equal choice() on p exactly; this runs
    the search over many workgroups, each of which reads cdf[-1]."""
    rng = np.random.default_rng(nstates + 17)
    p = rng.random(nstates) ** 2
    p = (p / p.sum()).astype(realtype)
    nshots = 40000
    backend.set_seed(4321)
    got = backend.to_numpy(backend.sample_shots((p * realtype(scale)).astype(realtype), nshots))
    want = np.random.RandomState(4321).choice(nstates, size=nshots, p=p.astype(np.float64))
    np.testing.assert_array_equal(got, want)


def test_sample_shots_stream_continues_like_numpy(backend):
    p = np.array([0.1, 0.2, 0.3, 0.4])
    backend.set_seed(7)
    a = backend.to_numpy(backend.sample_shots(p, 100))
    b = backend.to_numpy(backend.sample_shots(p, 50))
    rs = np.random.RandomState(7)
    np.testing.assert_array_equal(a, rs.choice(4, size=100, p=p))
    np.testing.assert_array_equal(b, rs.choice(4, size=50, p=p))
    assert backend.to_numpy(backend.sample_shots(p, 0)).shape == (0,)


def test_sample_frequencies_below_threshold_match_numpy(backend):
    p = np.zeros(32)
    p[[1, 5, 6, 17, 31]] = [0.1, 0.3, 0.2, 0.25, 0.15]
    backend.set_seed(99)
    got = backend.sample_frequencies(p, 20000)
    samples = np.random.RandomState(99).choice(32, size=20000, p=p)
    res, counts = np.unique(samples, return_counts=True)
    assert dict(got) == {int(k): int(v) for k, v in zip(res, counts)}


def test_calculate_frequencies_device_and_host_paths(backend):
    rng = np.random.default_rng(3)
    samples = rng.integers(0, 300, size=10000)
    want = dict(zip(*np.unique(samples, return_counts=True)))
    want = {int(k): int(v) for k, v in want.items()}
    dev = backend.calculate_frequencies(backend.cast(samples, dtype="int64"))
    host = backend.calculate_frequencies(samples)
    assert dict(dev) == want and dict(host) == want


def test_count_samples_rejects_out_of_range(backend):
    import ctypes

    from qibojit_amd import _native

    samples = backend.cast(np.array([0, 3, 7, -1, 2, 9]), dtype="int64")
    freqs = backend.cast(np.zeros(9, dtype=np.int64), dtype="int64")
    _native.call("qj_count_samples", ctypes.c_void_p(samples.data_ptr()), 6, ctypes.c_void_p(freqs.data_ptr()), 8,
                 ctypes.c_void_p(freqs.data_ptr() + 8 * 8), backend._stream())
    out = backend.to_numpy(freqs)
    np.testing.assert_array_equal(out[:8], [1, 0, 1, 1, 0, 0, 0, 1])
    assert out[8] == 2  # -1 and 9
    with pytest.raises(ValueError):
        backend._count_samples(samples, 8)
