"""NumPy restatement of the reference's stabilizer-tableau operations.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline; see oracle/__init__.py).

The tableau is the reference's unpacked symplectic matrix: a uint8 array of
shape (2n+1, 2n+1) holding 0/1, rows 0..n-1 the destabilizers, n..2n-1 the
stabilizers and 2n a scratch row; columns 0..n-1 the x bits, n..2n-1 the z
bits and 2n the sign bit r.  A row (x, z, r) stands for the Pauli string
(-1)^r * prod_k P(x_k, z_k) with P(0,0)=I, P(1,0)=X, P(1,1)=Y, P(0,1)=Z.

Every gate rewrites each row from the row's own bits: the update of the sign
is evaluated on the bits before the gate (src/qibojit/backends/
clifford_operations_cpu.py:18-248).  The reference applies the rules to 0/1
bytes with ``~``; every ``~`` there sits inside an AND with an un-negated
bit, so it acts as ``1 ^ v`` and the rules are restated on 0/1 values here.

How each function is pinned (tests/test_clifford.py):
  * the 15 gate rules: against U P U^dagger of the gate's textbook unitary
    on every row of random tableaux (brute force on 2 and 3 qubits);
  * ``rowsum``: the new sign against the matrix product P_i @ P_h of
    commuting rows (brute force), the determined mode against its definition
    on a zeroed target row;
  * ``random_outcome`` / ``determined_outcome``: stabilizers of the tableau
    built from |0...0> by random Clifford circuits against the state vector
    of the same circuit (every stabilizer fixes the state, a random outcome
    has probability 1/2, the collapsed state is fixed by the new
    stabilizers); for a determined outcome, the selected stabilizers
    multiply to +-Z_q, whose sign is the state's certain outcome, and the
    XOR rule equals the XOR of their signs (it misses the certain outcome on
    a recorded 3-qubit circuit, as the reference does);
  * ``determined_outcome_exact``: equal to the certain outcome on every
    determined case of those circuits (150 more at 2..6 qubits), including
    the case where the XOR rule misses it.
"""

import numpy as np


def _cols(sm, n):
    return sm[:, -1], sm[:, :n], sm[:, n:-1]


def _not(v):
    return 1 ^ v


# ---- one-qubit rules (clifford_operations_cpu.py:18-162) ----

def H(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= xq & zq
    x[:, q], z[:, q] = zq, xq
    return sm


def S(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= xq & zq
    z[:, q] = zq ^ xq
    return sm


def Z(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= (xq & zq) ^ (xq & (zq ^ xq))
    return sm


def X(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= (zq & (zq ^ xq)) ^ (zq & xq)
    return sm


def Y(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= (zq & (zq ^ xq)) ^ (xq & (zq ^ xq))
    return sm


def SX(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= zq & (zq ^ xq)
    x[:, q] = zq ^ xq
    return sm


def SDG(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= xq & (zq ^ xq)
    z[:, q] = zq ^ xq
    return sm


def SXDG(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= zq & xq
    x[:, q] = zq ^ xq
    return sm


def RY_pi(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= xq & (zq ^ xq)
    x[:, q], z[:, q] = zq, xq
    return sm


def RY_3pi_2(sm, q, n):
    r, x, z = _cols(sm, n)
    xq, zq = x[:, q].copy(), z[:, q].copy()
    r ^= zq & (zq ^ xq)
    x[:, q], z[:, q] = zq, xq
    return sm


# ---- two-qubit rules (clifford_operations_cpu.py:31-60, 165-248) ----

def _two(sm, c, t, n):
    r, x, z = _cols(sm, n)
    return r, x, z, x[:, c].copy(), z[:, c].copy(), x[:, t].copy(), z[:, t].copy()


def CNOT(sm, c, t, n):
    r, x, z, xc, zc, xt, zt = _two(sm, c, t, n)
    r ^= (xc & zt) & (xt ^ _not(zc))
    x[:, t] = xt ^ xc
    z[:, c] = zc ^ zt
    return sm


def CZ(sm, c, t, n):
    r, x, z, xc, zc, xt, zt = _two(sm, c, t, n)
    r ^= (xt & zt) ^ (xc & xt & (zt ^ _not(zc))) ^ (xt & (zt ^ xc))
    z[:, c] = xt ^ zc
    z[:, t] = zt ^ xc
    return sm


def SWAP(sm, c, t, n):
    r, x, z, xc, zc, xt, zt = _two(sm, c, t, n)
    r ^= ((xc & zt & (xt ^ _not(zc)))
          ^ ((xt ^ xc) & (zt ^ zc) & (zt ^ _not(xc)))
          ^ (xt & zc & (xc ^ xt ^ zc ^ _not(zt))))
    x[:, c], x[:, t] = xt, xc
    z[:, c], z[:, t] = zt, zc
    return sm


def iSWAP(sm, c, t, n):
    r, x, z, xc, zc, xt, zt = _two(sm, c, t, n)
    r ^= ((xt & zt) ^ (xc & zc) ^ (xc & (zc ^ xc))
          ^ ((zc ^ xc) & (zt ^ xt) & (xt ^ _not(xc)))
          ^ ((xt ^ zc ^ xc) & (xt ^ zt ^ xc) & (xt ^ zt ^ xc ^ _not(zc)))
          ^ (xc & (xt ^ xc ^ zc)))
    z[:, c] = xt ^ zt ^ xc
    z[:, t] = xt ^ zc ^ xc
    x[:, c], x[:, t] = xt, xc
    return sm


def CY(sm, c, t, n):
    r, x, z, xc, zc, xt, zt = _two(sm, c, t, n)
    r ^= ((xt & (zt ^ xt))
          ^ (xc & (xt ^ zt) & (zc ^ _not(xt)))
          ^ ((xt ^ xc) & (zt ^ xt)))
    x[:, t] = xc ^ xt
    z[:, c] = zc ^ zt ^ xt
    z[:, t] = zt ^ xc
    return sm


ONE_QUBIT = {"H": H, "S": S, "Z": Z, "X": X, "Y": Y, "SX": SX, "SDG": SDG, "SXDG": SXDG,
             "RY_pi": RY_pi, "RY_3pi_2": RY_3pi_2}
TWO_QUBIT = {"CNOT": CNOT, "CZ": CZ, "SWAP": SWAP, "iSWAP": iSWAP, "CY": CY}


def apply(sm, name, qubits, n):
    """Applies the rule ``name`` on ``qubits`` (a tuple of 1 or 2 ints)."""
    fn = ONE_QUBIT.get(name) or TWO_QUBIT[name]
    return fn(sm, *qubits, n)


# ---- rowsum and measurement (clifford_operations_cpu.py:337-385,
#      clifford_operations_gpu.py:606-640) ----

def _exponent(xi, zi, xh, zh):
    """Per-column exponent of i in P_i P_h (clifford_operations_cpu.py:361-365)."""
    xi, zi, xh, zh = (v.astype(np.int64) for v in (xi, zi, xh, zh))
    return 2 * (xi * xh * (zh - zi) + zi * zh * (xi - xh)) - xi * zh + xh * zi


def rowsum(sm, h, i, n, determined=False):
    """Row h[j] <- P_i[j] P_h[j] for every j, all read from the state before
    the call; ``determined`` instead XORs every product into row h[0]
    (clifford_operations_cpu.py:337-385)."""
    h = np.asarray(h, dtype=np.int64)
    i = np.asarray(i, dtype=np.int64)
    xi, zi = sm[i, :n].copy(), sm[i, n:-1].copy()
    xh, zh = sm[h, :n].copy(), sm[h, n:-1].copy()
    ri, rh = sm[i, -1].astype(np.int64), sm[h, -1].astype(np.int64)
    e = _exponent(xi, zi, xh, zh).sum(axis=1)
    rnew = ((2 * rh + 2 * ri + e) % 4 != 0).astype(np.uint8)
    if determined:
        acc_r = np.bitwise_xor.reduce(rnew) if len(rnew) else 0
        sm[h[0], -1] ^= acc_r
        sm[h[0], :n] ^= np.bitwise_xor.reduce(xi ^ xh, axis=0)
        sm[h[0], n:-1] ^= np.bitwise_xor.reduce(zi ^ zh, axis=0)
    else:
        sm[h, -1] = rnew
        sm[h, :n] = xi ^ xh
        sm[h, n:-1] = zi ^ zh
    return sm


def first_stabilizer(sm, q, n):
    """First stabilizer row (0-based among the stabilizers) with x_q = 1, or -1."""
    nz = np.nonzero(sm[n:-1, q])[0]
    return int(nz[0]) if len(nz) else -1


def random_outcome(sm, p, q, n, outcome):
    """Measurement of qubit q when stabilizer p anticommutes with Z_q
    (clifford_operations_gpu.py:606-625); ``outcome`` is the drawn bit."""
    p = int(p) + n
    h = [row for row in np.nonzero(sm[:-1, q])[0] if row != p]
    if h:
        rowsum(sm, h, [p] * len(h), n, False)
    sm[p - n, :] = sm[p, :]
    sm[p, :] = 0
    sm[p, -1] = outcome
    sm[p, n + q] = 1
    return sm, int(outcome)


def determined_outcome(sm, q, n):
    """Measurement of qubit q when every stabilizer commutes with Z_q
    (clifford_operations_gpu.py:628-640, rowsum in its determined mode)."""
    sm[-1, :] = 0
    idx = np.nonzero(sm[:n, q])[0] + n
    if len(idx) == 0:
        return sm, int(sm[-1, -1])
    rowsum(sm, [2 * n] * len(idx), idx, n, True)
    return sm, int(sm[-1, -1])


def determined_outcome_exact(sm, q, n):
    """Determined outcome with the phases of the product accumulated (the
    sequential Aaronson-Gottesman rowsum): the selected stabilizers S (rows
    n+i with destabilizer bit x_q(i) = 1, in row order) multiply to
    (-1)^(sum r_j + sum_{j<l} z_j.x_l) i^(sum |x_j & z_j|) Z_q, with Y = iXZ
    per qubit.  Not the reference's rule (which XORs the signs); the scratch
    row is left as ``determined_outcome`` leaves it, its sign set to the
    exact outcome."""
    sm, _ = determined_outcome(sm, q, n)
    rows = np.nonzero(sm[:n, q])[0] + n
    x = sm[rows, :n].astype(np.int64)
    z = sm[rows, n:-1].astype(np.int64)
    before = np.cumsum(z, axis=0) - z  # row l: sum of z_j over the rows j < l
    cross = int((before * x).sum())
    ys = int((x & z).sum())
    out = (int(sm[rows, -1].sum()) + cross + ys // 2) & 1
    sm[-1, -1] = out
    return sm, out


def zero_tableau(n):
    """Tableau of |0...0>: destabilizer k = X_k, stabilizer k = Z_k."""
    sm = np.zeros((2 * n + 1, 2 * n + 1), dtype=np.uint8)
    for k in range(n):
        sm[k, k] = 1
        sm[n + k, n + k] = 1
    return sm
