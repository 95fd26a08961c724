"""Kernel-facing gate encodings (src/qibojit/backends/matrices.py:7-70).

They are not always the mathematical matrix of the gate: a controlled gate
is passed as its target block (the controls select the amplitudes), U1/CU1
as the scalar phase e^{i theta} that the z_pow kernel multiplies by, fSim and
GeneralizedfSim as the 5-vector [u00, u01, u10, u11, e^{-i phi}] the fsim
kernel reads, DEUTSCH as i * RX(2 theta).  Matrices stay on the host: the
HIP library copies them into kernel arguments or its staging ring.
"""

from functools import cached_property

import numpy as np

from qibojit_amd._qibo import NumpyMatrices


class CustomMatrices(NumpyMatrices):
    """Matrices used by the custom operators."""

    def CRX(self, theta):
        return self.RX(theta)

    def CRY(self, theta):
        return self.RY(theta)

    def CRZ(self, theta):
        return self.RZ(theta)

    def CU2(self, phi, lam):
        return self.U2(phi, lam)

    def CU3(self, theta, phi, lam):
        return self.U3(theta, phi, lam)

    def U1(self, theta):
        return np.asarray(np.exp(1j * theta), dtype=self.dtype)

    def CU1(self, theta):
        return self.U1(theta)

    @cached_property
    def CZ(self):
        return self.Z

    @cached_property
    def CCZ(self):
        return self.Z

    @cached_property
    def CY(self):
        return self.Y

    @cached_property
    def CH(self):
        return self.H

    @cached_property
    def CSX(self):
        return self.SX

    @cached_property
    def CSXDG(self):
        return self.SXDG

    def DEUTSCH(self, theta):
        return 1j * self.RX(2 * theta)

    def fSim(self, theta, phi):
        cost = np.cos(theta) + 0j
        isint = -1j * np.sin(theta)
        phase = np.exp(-1j * phi)
        return np.array([cost, isint, isint, cost, phase], dtype=self.dtype)

    def GeneralizedfSim(self, u, phi):
        u = np.asarray(u)
        phase = np.exp(-1j * phi)
        return np.array([u[0, 0], u[0, 1], u[1, 0], u[1, 1], phase], dtype=self.dtype)
