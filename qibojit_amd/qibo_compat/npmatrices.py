"""Dense gate matrices in qibo's conventions (qibo.backends.npmatrices).

Row-major, big-endian in the order the gate lists its qubits (controls
first for the named controlled gates, e.g. CNOT(control, target)).  These are
the textbook forms qibo documents for each gate.
"""

import cmath
import math
from functools import cached_property

import numpy as np

# float covers numpy.float64; float32 / float16 angles keep the NumPy forms,
# which evaluate in their own precision
_REAL_SCALARS = (float, int, np.integer)


def _scalars(*xs):
    """Whether every angle is a real double-precision or integer scalar:
    those take the scalar `math` / `cmath` forms below (one array built per
    matrix, a few microseconds less per gate than NumPy scalar arithmetic;
    the same formulas, equal to within an ulp)."""
    return all(isinstance(x, _REAL_SCALARS) and not isinstance(x, bool) for x in xs)


class NumpyMatrices:
    """Matrices of qibo's gates, cast to ``dtype``."""

    def __init__(self, dtype):
        self.dtype = dtype
        self.np = np

    def _cast(self, x, dtype):
        return np.array(x, dtype=dtype)

    # ---- one qubit ----
    def I(self, n=2):  # noqa: E743
        return np.eye(n, dtype=self.dtype)

    @cached_property
    def H(self):
        return self._cast(np.array([[1, 1], [1, -1]]) / np.sqrt(2), dtype=self.dtype)

    @cached_property
    def X(self):
        return self._cast([[0, 1], [1, 0]], dtype=self.dtype)

    @cached_property
    def Y(self):
        return self._cast([[0, -1j], [1j, 0]], dtype=self.dtype)

    @cached_property
    def Z(self):
        return self._cast([[1, 0], [0, -1]], dtype=self.dtype)

    @cached_property
    def SX(self):
        return self._cast(0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]), dtype=self.dtype)

    @cached_property
    def SXDG(self):
        return self._cast(0.5 * np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]]), dtype=self.dtype)

    @cached_property
    def S(self):
        return self._cast([[1, 0], [0, 1j]], dtype=self.dtype)

    @cached_property
    def SDG(self):
        return self._cast([[1, 0], [0, -1j]], dtype=self.dtype)

    @cached_property
    def T(self):
        return self._cast([[1, 0], [0, np.exp(1j * np.pi / 4)]], dtype=self.dtype)

    @cached_property
    def TDG(self):
        return self._cast([[1, 0], [0, np.exp(-1j * np.pi / 4)]], dtype=self.dtype)

    def RX(self, theta):
        if _scalars(theta):
            cos, isin = math.cos(theta / 2), -1j * math.sin(theta / 2)
            return np.array([[cos, isin], [isin, cos]], dtype=self.dtype)
        cos, isin = np.cos(theta / 2), -1j * np.sin(theta / 2)
        return self._cast([[cos, isin], [isin, cos]], dtype=self.dtype)

    def RY(self, theta):
        if _scalars(theta):
            cos, sin = math.cos(theta / 2), math.sin(theta / 2)
            return np.array([[cos, -sin], [sin, cos]], dtype=self.dtype)
        cos, sin = np.cos(theta / 2), np.sin(theta / 2)
        return self._cast([[cos, -sin], [sin, cos]], dtype=self.dtype)

    def RZ(self, theta):
        if _scalars(theta):
            phase = cmath.exp(0.5j * theta)
            return np.array([[phase.conjugate(), 0], [0, phase]], dtype=self.dtype)
        phase = np.exp(0.5j * theta)
        return self._cast([[np.conj(phase), 0], [0, phase]], dtype=self.dtype)

    def U1(self, theta):
        if _scalars(theta):
            return np.array([[1, 0], [0, cmath.exp(1j * theta)]], dtype=self.dtype)
        return self._cast([[1, 0], [0, np.exp(1j * theta)]], dtype=self.dtype)

    def U2(self, phi, lam):
        eplus = np.exp(1j * (phi + lam) / 2.0)
        eminus = np.exp(1j * (phi - lam) / 2.0)
        return self._cast(
            np.array([[np.conj(eplus), -np.conj(eminus)], [eminus, eplus]]) / np.sqrt(2),
            dtype=self.dtype,
        )

    def U3(self, theta, phi, lam):
        if _scalars(theta, phi, lam):
            cost, sint = math.cos(theta / 2), math.sin(theta / 2)
            eplus = cmath.exp(1j * (phi + lam) / 2.0)
            eminus = cmath.exp(1j * (phi - lam) / 2.0)
            return np.array(
                [[eplus.conjugate() * cost, -eminus.conjugate() * sint], [eminus * sint, eplus * cost]],
                dtype=self.dtype,
            )
        cost, sint = np.cos(theta / 2), np.sin(theta / 2)
        eplus = np.exp(1j * (phi + lam) / 2.0)
        eminus = np.exp(1j * (phi - lam) / 2.0)
        return self._cast(
            [[np.conj(eplus) * cost, -np.conj(eminus) * sint], [eminus * sint, eplus * cost]],
            dtype=self.dtype,
        )

    # ---- two qubits ----
    @staticmethod
    def _controlled(u):
        u = np.asarray(u)
        d = u.shape[0]
        m = np.eye(2 * d, dtype=u.dtype)
        m[d:, d:] = u
        return m

    @cached_property
    def CNOT(self):
        return self._cast(self._controlled(self.X), dtype=self.dtype)

    @cached_property
    def CY(self):
        return self._cast(self._controlled(self.Y), dtype=self.dtype)

    @cached_property
    def CZ(self):
        return self._cast(self._controlled(self.Z), dtype=self.dtype)

    @cached_property
    def CSX(self):
        return self._cast(self._controlled(self.SX), dtype=self.dtype)

    @cached_property
    def CSXDG(self):
        return self._cast(self._controlled(self.SXDG), dtype=self.dtype)

    @cached_property
    def CH(self):
        return self._cast(self._controlled(self.H), dtype=self.dtype)

    def CRX(self, theta):
        return self._cast(self._controlled(self.RX(theta)), dtype=self.dtype)

    def CRY(self, theta):
        return self._cast(self._controlled(self.RY(theta)), dtype=self.dtype)

    def CRZ(self, theta):
        return self._cast(self._controlled(self.RZ(theta)), dtype=self.dtype)

    def CU1(self, theta):
        return self._cast(self._controlled(self.U1(theta)), dtype=self.dtype)

    def CU2(self, phi, lam):
        return self._cast(self._controlled(self.U2(phi, lam)), dtype=self.dtype)

    def CU3(self, theta, phi, lam):
        return self._cast(self._controlled(self.U3(theta, phi, lam)), dtype=self.dtype)

    @cached_property
    def SWAP(self):
        return self._cast(
            [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=self.dtype
        )

    @cached_property
    def iSWAP(self):
        return self._cast(
            [[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=self.dtype
        )

    @cached_property
    def FSWAP(self):
        return self._cast(
            [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, -1]], dtype=self.dtype
        )

    def fSim(self, theta, phi):
        cost, isint = np.cos(theta), -1j * np.sin(theta)
        return self._cast(
            [[1, 0, 0, 0], [0, cost, isint, 0], [0, isint, cost, 0], [0, 0, 0, np.exp(-1j * phi)]],
            dtype=self.dtype,
        )

    def GeneralizedfSim(self, u, phi):
        m = np.eye(4, dtype=complex)
        m[1:3, 1:3] = np.asarray(u)
        m[3, 3] = np.exp(-1j * phi)
        return self._cast(m, dtype=self.dtype)

    def RXX(self, theta):
        cos, isin = np.cos(theta / 2), -1j * np.sin(theta / 2)
        return self._cast(
            [[cos, 0, 0, isin], [0, cos, isin, 0], [0, isin, cos, 0], [isin, 0, 0, cos]],
            dtype=self.dtype,
        )

    def RYY(self, theta):
        cos, isin = np.cos(theta / 2), -1j * np.sin(theta / 2)
        return self._cast(
            [[cos, 0, 0, -isin], [0, cos, isin, 0], [0, isin, cos, 0], [-isin, 0, 0, cos]],
            dtype=self.dtype,
        )

    def RZZ(self, theta):
        a, b = np.exp(-0.5j * theta), np.exp(0.5j * theta)
        return self._cast(np.diag([a, b, b, a]), dtype=self.dtype)

    # ---- three qubits ----
    @cached_property
    def TOFFOLI(self):
        return self._cast(self._controlled(self._controlled(self.X)), dtype=self.dtype)

    @cached_property
    def CCZ(self):
        return self._cast(self._controlled(self._controlled(self.Z)), dtype=self.dtype)

    def DEUTSCH(self, theta):
        m = np.eye(8, dtype=complex)
        m[6:, 6:] = 1j * np.asarray(self.RX(2 * theta), dtype=complex)
        return self._cast(m, dtype=self.dtype)

    # ---- generic ----
    def Unitary(self, u):
        return self._cast(u, dtype=self.dtype)

    def FanOut(self, *q):
        # CNOT from q[0] onto each of q[1:], over len(q) qubits.
        n = len(q)
        dim = 2**n
        m = np.zeros((dim, dim), dtype=self.dtype)
        ntargets = n - 1
        for i in range(dim):
            j = i ^ ((1 << ntargets) - 1) if (i >> ntargets) & 1 else i
            m[j, i] = 1
        return m
