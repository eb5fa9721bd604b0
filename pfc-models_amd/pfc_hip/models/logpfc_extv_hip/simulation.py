"""log PFC with the M-shifted split and an external potential on MI355X: the
drop-in sibling of the reference's models/logpfc_extv/simulation.py.

    d_t phi = Gamma lap (L_mu phi + ln phi + g phi^2 + phi^3) + lap V,
    phi_hat' = E phi_hat + P (Gamma lap N_hat + lap V_hat)   (:557-577)

with the log clamped at minPhi (:524-539), L_mu shifted by M and the
nonlinearity by -M phi, M re-fitted when phi leaves its window by more than
10 % (OptimizeM :332-373, ValidatePhiRange :375-405), and the k = 0 mode
pinned (:577).  The library carries the potential as a constant source term
P lap V_hat per step (pfc_set_external_potential); the free energy excludes
it, as the reference's does.  The deferred predictor-corrector of the
reference (model.md, with its `~~small_z` index slip) is the `scheme="etdrk2"`
option here, with the Cox-Matthews phi_2 including its 1/dt.
"""
from __future__ import annotations

import numpy as np

from pfc_hip.discretization import Build2DGeometry, BuildDerivativeKernels, kernel_etd, kernel_local  # noqa: F401
from pfc_hip.models._base import ModelParameters
from pfc_hip.models.logpfc_hip.simulation import GetLinearKernel_mu, GetLinearKernel_phi  # noqa: F401
from pfc_hip.models.logpfc_hip.simulation import Simulation as _LogSimulation

from .external_potentials import ExternalPotential


def BuildLogPFCModel_withVacancyForce_Parameters(Gamma=1.0, beta=1.0, q0=1.0, epsilon=1.0, g=1.0, dt=0.01,
                                                 minPhi=1e-6, meanDensity=None, smallZTol=1e-8, V0=None,
                                                 potWidth=None):
    """models/logpfc_extv/simulation.py:69-97 (defaults included: smallZTol 1e-8)."""
    model = ModelParameters()
    model.Gamma = Gamma
    model.beta = beta
    model.q0 = q0
    model.epsilon = epsilon
    model.g = g
    model.dt = dt
    model.minPhi = minPhi
    model.meanDensity = meanDensity
    model.smallZTol = smallZTol
    model.V0 = V0
    model.potWidth = potWidth
    model.minLog = np.log(minPhi)
    return model


def BuildGeometry(Nx, Ny, Lx, Ly):
    """Build2DGeometry plus the triangular-lattice constants a = 4 pi / sqrt 3
    and h = 4 pi that this model reads (:487-497)."""
    geometry = Build2DGeometry(Nx, Ny, Lx, Ly)
    geometry.a = 4 * np.pi / 3 ** 0.5
    geometry.h = 4 * np.pi
    return geometry


class Simulation(_LogSimulation):
    """The logPFC simulation with the stabilised split always on and an
    external potential (`_externalV`, the reference's attribute)."""

    def __init__(self, modelParameters, geometry, **kw):
        super().__init__(modelParameters, geometry, **kw)
        self._externalV = ExternalPotential(geometry, self)

    def _CalcLapV(self):
        """lap V in k-space and real space, on the host (:273-283)."""
        ev = self._externalV
        lap = BuildDerivativeKernels(self.Geometry).d2_dlap
        ev.lap_V_hat = lap * ev.V_hat
        ev.lap_V = np.fft.irfft2(ev.lap_V_hat, s=ev.V.shape)

    def GetForce(self):
        """F = -grad V, shape (2, Ny, Nx) (:453-468)."""
        dk = BuildDerivativeKernels(self.Geometry)
        V_hat = self._externalV.V_hat
        s = self._externalV.V.shape
        return -np.stack([np.fft.irfft2(dk.d_dx * V_hat, s=s), np.fft.irfft2(dk.d_dy * V_hat, s=s)])

    def CalculateChemicalPotential(self):
        """mu = dF/dphi of the current field as a host array (:419-425); the
        potential is not part of mu.  State.local_mu is filled as well."""
        self._Batch1_CalcNonlinMu()
        self._Batch2_CalcMu_Terms()
        return self.State.mu


def BuildHipGPU_Simulation(modelParameters, geometry, *, precision=64, scheme="etd1"):
    """Counterpart of BuildCupyGPU_Simulation (:500-521): clamped log, M-shifted split."""
    return Simulation(modelParameters, geometry, precision=precision, scheme=scheme, pin_mean=True, clamp_log=True,
                      stabilise=True)
