"""Textbook gate matrices for the oracle (complex128), written independently
of the product's matrix module.  Conventions: qubit order as listed, first
qubit = most significant bit; controlled gates list control(s) first.
TEST INFRASTRUCTURE ONLY.
"""

import numpy as np

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
S = np.diag([1, 1j])
SDG = np.diag([1, -1j])
T = np.diag([1, np.exp(1j * np.pi / 4)])
TDG = np.diag([1, np.exp(-1j * np.pi / 4)])
# sqrt(X) = (1+i)/2 I + (1-i)/2 X
SX = 0.5 * (1 + 1j) * I2 + 0.5 * (1 - 1j) * X
SXDG = SX.conj().T


def expm_pauli(theta, p):
    """exp(-i theta/2 P) for a Pauli P (P^2 = I)."""
    return np.cos(theta / 2) * np.eye(p.shape[0]) - 1j * np.sin(theta / 2) * p


def RX(theta):
    return expm_pauli(theta, X)


def RY(theta):
    return expm_pauli(theta, Y)


def RZ(theta):
    return expm_pauli(theta, Z)


def U1(theta):
    return np.diag([1, np.exp(1j * theta)])


def U3(theta, phi, lam):
    # e^{i(phi+lam)/2} RZ(phi) RY(theta) RZ(lam), the global phase chosen so
    # that the determinant is 1.
    return RZ(phi) @ RY(theta) @ RZ(lam)


def U2(phi, lam):
    return U3(np.pi / 2, phi, lam)


def controlled(u, ncontrols=1):
    u = np.asarray(u, dtype=complex)
    for _ in range(ncontrols):
        d = u.shape[0]
        m = np.eye(2 * d, dtype=complex)
        m[d:, d:] = u
        u = m
    return u


SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)


def fSim(theta, phi):
    m = np.eye(4, dtype=complex)
    m[1:3, 1:3] = [[np.cos(theta), -1j * np.sin(theta)], [-1j * np.sin(theta), np.cos(theta)]]
    m[3, 3] = np.exp(-1j * phi)
    return m


def GeneralizedfSim(u, phi):
    m = np.eye(4, dtype=complex)
    m[1:3, 1:3] = np.asarray(u)
    m[3, 3] = np.exp(-1j * phi)
    return m


def DEUTSCH(theta):
    return controlled(1j * RX(2 * theta), 2)


def haar_unitary(dim, rng):
    """Haar-random unitary: QR of a complex Ginibre matrix with the phases of
    R's diagonal moved into Q (Mezzadri 2007)."""
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))


def random_state(nqubits, rng, dtype=np.complex128):
    """Normalised complex Gaussian state."""
    n = 2**nqubits
    psi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    psi /= np.linalg.norm(psi)
    return psi.astype(dtype)


def random_density_matrix(nqubits, rng, dtype=np.complex128):
    """Normalised Hermitian positive matrix G G^dagger / tr."""
    d = 2**nqubits
    g = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    rho = g @ g.conj().T
    rho /= np.trace(rho)
    return rho.astype(dtype)
