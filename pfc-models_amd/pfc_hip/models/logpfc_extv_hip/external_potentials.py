"""External potential shapes of logpfc_extv (the reference's
models/logpfc_extv/external_potentials.py): a generalised Gaussian for atom
fits, a Gaussian dipole, a tanh well along x and a tanh double vacancy.

The potential V(x, y) is built on the host in FP64 and handed to the
library (pfc_set_external_potential), which keeps rfft2(V) and the update's
source term P lap V_hat on the device; `V` and `V_hat` here are host copies.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.optimize import curve_fit


@dataclass
class FitParameters:
    coords: tuple
    A0: float
    sigma: float
    b: float


@dataclass
class FitResult:
    inputCoords: tuple
    outputParameters: FitParameters
    chi2: float
    pcov: np.ndarray
    atomNeighborhood: np.ndarray
    X_Neighborhood: np.ndarray
    Y_Neighborhood: np.ndarray


def GeneralizedGaussian(xy_grid, x0, y0, A0, sigma0, b0):
    """b0 + A0 exp(-|r - r0|^2 / (2 sigma0^2)) (external_potentials.py:41-44)."""
    x, y = xy_grid
    return b0 + A0 * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * sigma0 ** 2))


class ExternalPotential:
    """V(x, y) and V_hat of one simulation (external_potentials.py:51-131);
    every setter uploads V to the simulation's device plan."""

    def __init__(self, geometry, sim=None):
        self.Geometry = geometry
        self._sim = sim
        shape = (geometry.Ny, geometry.Nx)
        self.X = np.asarray(geometry.X, dtype=np.float64)
        self.Y = np.asarray(geometry.Y, dtype=np.float64)
        self.V = np.zeros(shape)
        self.V_hat = np.fft.rfft2(self.V)
        self.lap_V = np.zeros(shape)
        self.lap_V_hat = np.fft.rfft2(self.lap_V)

    def _set(self, V):
        self.V = np.asarray(V, dtype=np.float64).reshape(self.X.shape)
        self.V_hat = np.fft.rfft2(self.V)
        if self._sim is not None:
            self._sim._plan.set_external_potential(self.V)

    def clear(self):
        """No potential (the update has no source term)."""
        self.V = np.zeros_like(self.V)
        self.V_hat = np.fft.rfft2(self.V)
        if self._sim is not None:
            self._sim._plan.set_external_potential(None)

    def SetGaussianDipolePotential(self, V0, X0, X1, sigma):
        """V0 [G(r - X0) - G(r - X1)], G = exp(-r^2 / (2 sigma^2)) (:109-119)."""
        (x0, y0), (x1, y1) = X0, X1
        s2 = 2 * sigma ** 2
        self._set(V0 * (np.exp(-((self.X - x0) ** 2 + (self.Y - y0) ** 2) / s2)
                        - np.exp(-((self.X - x1) ** 2 + (self.Y - y1) ** 2) / s2)))

    def SetPotential(self, V0, V_func):
        """V0 V_func, V_func an (Ny, Nx) array (:121-123)."""
        self._set(V0 * np.asarray(V_func, dtype=np.float64))

    def SetPotential_alongX(self, V0, width, x1, x2):
        """V0 tanh((|x - (x1 + x2)/2| - (x2 - x1)/2) / width): a well between x1 and x2 (:125-127)."""
        self._set(V0 * np.tanh((np.abs(self.X - (x1 + x2) / 2) - (x2 - x1) / 2) / width))

    def SetDoubleVacancyAlongX_TanhPotential(self, x1, x2, y1, y2, width_x, width_y, V0):
        """The tanh double-vacancy well of :129-131: 1 outside, 1 - V0 inside
        the (x1..x2) x (y1..y2) box, edges of widths width_x, width_y."""
        tx = (np.tanh((np.abs(self.X - (x1 + x2) / 2) - (x2 - x1) / 2) / width_x) - 1) / 2
        ty = (1 - np.tanh((np.abs(self.Y - (y1 + y2) / 2) - (y2 - y1) / 2) / width_y)) / 2
        self._set(V0 * tx * ty + 1)

    def FitAtomToGaussian(self, phi, x0, y0):
        """Fit the atom near (x0, y0) in `phi` (a host field) to a generalised
        Gaussian over a one-lattice-spacing window (:62-107)."""
        g = self.Geometry
        a = g.a
        lo_x, hi_x = g.x_to_ix(x0 - a / 2), g.x_to_ix(x0 + a / 2)
        lo_y, hi_y = g.y_to_iy(y0 - a / 2), g.y_to_iy(y0 + a / 2)
        hood = np.asarray(phi)[lo_y:hi_y + 1, lo_x:hi_x + 1]
        Xn, Yn = self.X[lo_y:hi_y + 1, lo_x:hi_x + 1], self.Y[lo_y:hi_y + 1, lo_x:hi_x + 1]
        grid = (Xn.ravel(), Yn.ravel())
        p0 = [x0, y0, hood.max() - hood.min(), a / 4, hood.min()]
        popt, pcov = curve_fit(GeneralizedGaussian, grid, hood.ravel(), p0=p0)
        chi2 = float(np.sum((GeneralizedGaussian(grid, *popt) - hood.ravel()) ** 2))
        xr, yr, A0, sigma0, b0 = popt
        return FitResult(inputCoords=(x0, y0), outputParameters=FitParameters(coords=(xr, yr), A0=A0, sigma=sigma0,
                                                                               b=b0),
                         chi2=chi2, pcov=pcov, atomNeighborhood=hood, X_Neighborhood=Xn, Y_Neighborhood=Yn)
