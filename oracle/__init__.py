"""CPU oracle of the state-vector path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything here, and only to check or time against it.
The product (``qibojit_amd``) never imports, loads or calls this package.

Contents
--------
einsum_ref
    Closed forms: gates applied by ``numpy.tensordot`` on the state reshaped
    to ``(2,) * n`` (what qibo's NumpyBackend does and what the reference's
    tests compare against), density matrices, collapse by slicing, marginal
    probabilities, TFIM energy, and brute-force Kronecker products.
gate_matrices
    Textbook gate matrices, written independently of the product's.
numba_port.c / port.py
    C/OpenMP restatement of qibojit's numba kernels
    (src/qibojit/custom_operators/gates.py, ops.py:5-108) with a Python
    dispatcher restating NumbaBackend (src/qibojit/backends/cpu.py:369-635),
    including the NumPy-legacy MT19937 Metropolis sampler.  It is the
    NumbaBackend stand-in for parity and the timed CPU baseline.

How each part is pinned (so that a plausible slip fails a check):
  * einsum_ref gates: against Kronecker-product brute force on small n,
    unitarity/inverse round trips, and identities between gates.
  * gate_matrices: identities U3(t,-pi/2,pi/2)=RX(t), U3(t,0,0)=RY(t),
    U1 = e^{i l/2} RZ(l), H X H = Z, SX^2 = X, fSim(pi/2,0) = iSWAP^dagger
    pattern, and unitarity.
  * port kernels: against einsum_ref on the reference's own case lists.
  * Metropolis sampler: the reference's golden vector
    (tests/test_ops.py:236-256) and numpy.random.RandomState streams.
  * collapse / probabilities: slicing closed forms (tests/test_ops.py:65-83).
  * fused blocks: fused == unfused execution.
  * clifford_ref: every rule against U P U^dagger of the gate's unitary,
    rowsum against Pauli products, outcomes against the state vector of
    the same Clifford circuit (tests/test_clifford.py).
"""
