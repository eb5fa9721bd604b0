"""log PFC, 2D spectral, MI355X (HIP kernels through libpfc_hip.so).

The drop-in sibling of the reference's models/logpfc/simulation.py and
models/logpfc_cpu/simulation.py.  The field is `phi`;

    f = phi beta/2 (q0^2 + lap)^2 phi + phi ln phi + eps/2 phi^2 + g/3 phi^3 + phi^4/4,
    d_t phi = Gamma lap (L_mu phi + ln phi + g phi^2 + phi^3).

Options beyond the reference logpfc (both documented there):
  clamp_log  ln(phi) -> ln(minPhi) where phi <= minPhi, the clamp form of
             logpfc_extv (simulation.py:536-539; CONTEXT.md:33-35);
  stabilise  the M-shifted linear/nonlinear split of logpfc_extv
             (GetLinearKernel_mu(M) :100-116, OptimizeM :332-373,
             ValidatePhiRange :375-405): +M on L_mu, -M phi on the
             nonlinearity, M re-fitted when the phi range drifts by > 10 %.
"""
from __future__ import annotations

import math

import numpy as np

from pfc_hip.backends.hip_rocm import npComplexType, npFloatType  # noqa: F401
from pfc_hip.discretization import (  # noqa: F401
    Build2DGeometry,
    BuildDerivativeKernels,
    Geometry,
    kernel_etd,
    kernel_local,
)
from pfc_hip.models._base import ModelParameters, Simulation as _Simulation


def BuildLogPFCModelParameters(Gamma=1.0, beta=1.0, q0=1.0, epsilon=1.0, g=1.0, dt=0.01, minPhi=1e-6,
                               meanDensity=None, smallZTol=2e-5):
    """models/logpfc/simulation.py:51-75 (defaults included)."""
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
    model.minLog = np.log(minPhi)
    return model


def GetLinearKernel_mu(modelParameters, DerivativeKernels, M=0):
    m = modelParameters
    return m.beta * (m.q0**2 + DerivativeKernels.d2_dlap)**2 + m.epsilon + M


def GetLinearKernel_phi(modelParameters, DerivativeKernels, M=0):
    return modelParameters.Gamma * DerivativeKernels.d2_dlap * GetLinearKernel_mu(modelParameters, DerivativeKernels, M)


class Simulation(_Simulation):
    _model_kind = "log"
    _field_name = "phi"

    def __init__(self, modelParameters, geometry, *, stabilise=False, **kw):
        host = stabilise == "host"
        super().__init__(modelParameters, geometry, stabilise=bool(stabilise), **kw)
        s = self.State
        s.phi_floor_tol = 0.1
        s.phi_ceiling_tol = 0.1
        self._host_shift = self.stabilise and host
        # stabilise=True: the library checks the phi range and refits M on the
        # device before every step (pfc_set_auto_shift), with no host round
        # trip; State.M / phi_min / phi_max then read the device's values.
        # stabilise="host": the reference's per-step host check (below).
        self._device_shift = self.stabilise and not host
        if self._device_shift:
            self._plan.set_auto_shift(True, s.phi_floor_tol, s.phi_ceiling_tol)
        else:
            s.M = 0.0
            s.phi_min = None
            s.phi_max = None

    @property
    def phi(self):
        return self.State.phi

    def _upload(self, value):
        super()._upload(value)
        # a new field refits M at the next step (on the device, the library
        # resets the fit itself)
        if not getattr(self, "_device_shift", False) and getattr(self.State, "phi_min", None) is not None:
            self.State.phi_min = self.State.phi_max = None

    def Get_local_phi(self):
        """Gaussian-smoothed density (w = 2 pi) as a host array; the reference's
        axes=(-2, 1) slip (models/logpfc/simulation.py:240) is not reproduced."""
        return self._local()

    # -- stabilised split (logpfc_extv) ---------------------------------------
    def OptimizeM(self):
        """M = (N(phi_max) - N(phi_min)) / (phi_max - phi_min), N = ln phi + g phi^2 + phi^3,
        phi_min replaced by minPhi when the field is not positive; rebuilds the
        device ETD tables (logpfc_extv/simulation.py:332-373).  With the
        device-managed shift, makes the next step refit."""
        if getattr(self, "_device_shift", False):
            s = self.State
            self._plan.set_auto_shift(True, s.phi_floor_tol, s.phi_ceiling_tol)
            return
        _, _, mn, mx = self._plan.reduce()
        g = float(self.Model.g)
        lo = mn if mn > 0 else float(self.Model.minPhi)
        n_hi = g * mx**2 + mx**3 + math.log(mx)
        n_lo = g * lo**2 + lo**3 + math.log(lo)
        M = (n_hi - n_lo) / (mx - lo)
        self._plan.set_shift(M)
        self.State.phi_min, self.State.phi_max, self.State.M = lo, mx, M

    def ValidatePhiRange(self):
        """Re-fit M when phi leaves the fitted window by more than the tolerances
        (logpfc_extv/simulation.py:375-405).  With the device-managed shift the
        library does this before every step, so there is nothing to do here."""
        if getattr(self, "_device_shift", False):
            return
        s = self.State
        if s.phi_min is None or s.phi_max is None:
            self.OptimizeM()
            return
        _, _, mn, mx = self._plan.reduce()
        floor_breached = (s.phi_min - mn) / s.phi_min > s.phi_floor_tol
        ceiling_breached = (mx - s.phi_max) / s.phi_max > s.phi_ceiling_tol
        if floor_breached or ceiling_breached:
            self.OptimizeM()


def BuildHipGPU_Simulation(modelParameters, geometry, *, precision=64, scheme="etd1", pin_mean=True,
                           clamp_log=None, stabilise=False):
    """Counterpart of BuildCupyGPU_Simulation (models/logpfc/simulation.py:265-281).
    FP32 bulk always clamps the log (the guard near phi = 0).  stabilise:
    False, True (the M-shifted split with the range check and refit on the
    device every step) or "host" (the same check on the host, as logpfc_extv
    does, with a synchronisation per step)."""
    if clamp_log is None:
        clamp_log = precision == 32 or stabilise
    return Simulation(modelParameters, geometry, precision=precision, scheme=scheme, pin_mean=pin_mean,
                      clamp_log=clamp_log, stabilise=stabilise)
