"""standard PFC, 2D spectral, MI355X (HIP kernels through libpfc_hip.so).

The drop-in sibling of the reference's models/stdpfc/simulation.py and
models/stdpfc_cpu/simulation.py: same builders and names.  The field is `n`;

    f = n beta/2 (q0^2 + lap)^2 n + eps/2 n^2 + g/3 n^3 + n^4/4,
    d_t n = Gamma lap (L_mu n + g n^2 + n^3),   L_mu = beta (q0^2 + lap)^2 + eps,

advanced by ETD1 (the reference's scheme) or ETDRK2 in Fourier space, in FP64
or FP32 bulk arithmetic with FP64-built tables and FP64 reductions.
"""
from __future__ import annotations

from pfc_hip.backends.hip_rocm import npComplexType, npFloatType  # noqa: F401
from pfc_hip.discretization import (  # noqa: F401  (re-exported, as the reference's module does)
    Build2DGeometry,
    BuildDerivativeKernels,
    Geometry,
    kernel_etd,
    kernel_local,
)
from pfc_hip.models._base import ModelParameters, Simulation as _Simulation


def BuildStdPFCModelParameters(Gamma=1.0, beta=1.0, q0=1.0, epsilon=1.0, g=1.0, dt=0.01, meanDensity=None,
                               smallZTol=2e-5):
    """models/stdpfc/simulation.py:60-80."""
    model = ModelParameters()
    model.Gamma = Gamma
    model.beta = beta
    model.q0 = q0
    model.epsilon = epsilon
    model.g = g
    model.dt = dt
    model.meanDensity = meanDensity
    model.smallZTol = smallZTol
    return model


def GetLinearKernel_mu(modelParameters, DerivativeKernels):
    """Host FP64 L_mu = beta (q0^2 + lap)^2 + eps (models/stdpfc/simulation.py:83-89)."""
    m = modelParameters
    return m.beta * (m.q0**2 + DerivativeKernels.d2_dlap)**2 + m.epsilon


def GetLinearKernel_n(modelParameters, DerivativeKernels):
    """Host FP64 L = Gamma lap L_mu (models/stdpfc/simulation.py:91-96)."""
    return modelParameters.Gamma * DerivativeKernels.d2_dlap * GetLinearKernel_mu(modelParameters, DerivativeKernels)


class Simulation(_Simulation):
    _model_kind = "std"
    _field_name = "n"

    @property
    def n(self):
        return self.State.n

    def Get_local_n(self):
        """Gaussian-smoothed density (w = 2 pi) as a host array (models/stdpfc/simulation.py:237-247)."""
        return self._local()


def BuildHipGPU_Simulation(modelParameters, geometry, *, precision=64, scheme="etd1", pin_mean=True):
    """Counterpart of BuildCupyGPU_Simulation (models/stdpfc/simulation.py:270-286)."""
    return Simulation(modelParameters, geometry, precision=precision, scheme=scheme, pin_mean=pin_mean)
