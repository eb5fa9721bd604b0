"""Host-side 2D periodic spectral geometry and Fourier-operator symbols.

The names and the contract are the reference substrate's
(discretization/spectral/geometry_2d.py:38-88, kernels.py:33-71), so a
caller written against them works unchanged with the `hip` models.  These
FP64 NumPy arrays describe the grid for callers (coordinates, k-grids, the
Gaussian symbol); the device ETD tables are built by libpfc_hip.so itself
from (Nx, Ny, Lx, Ly) and the model parameters.
"""
from __future__ import annotations

import numpy as np


class Geometry:
    """Nx, Ny, Lx, Ly, dx, dy, x, y, X, Y, kx, ky, KX, KY, K2 and the
    coordinate/index conversions (geometry_2d.py:38-67)."""

    def __init__(self):
        self.Nx = None
        self.Ny = None
        self.Lx = None
        self.Ly = None
        self.dx = None
        self.dy = None
        self.x = None
        self.y = None
        self.X = None
        self.Y = None
        self.periodic = (True, True)

    def px_to_x(self, px):
        return px * self.dx

    def py_to_y(self, py):
        return py * self.dy

    def x_to_px(self, x):
        return x / self.dx

    def y_to_py(self, y):
        return y / self.dy

    def x_to_ix(self, x):
        return int(x / self.dx)

    def y_to_iy(self, y):
        return int(y / self.dy)


def Build2DGeometry(Nx, Ny, Lx, Ly):
    """kx = 2 pi rfftfreq(Nx, dx), ky = 2 pi fftfreq(Ny, dy); meshgrids in (Ny, .) order."""
    g = Geometry()
    g.Nx, g.Ny, g.Lx, g.Ly = Nx, Ny, Lx, Ly
    g.dx = g.Lx / g.Nx
    g.dy = g.Ly / g.Ny
    g.x = np.linspace(0, g.Lx, g.Nx, endpoint=False, dtype=np.float64)
    g.y = np.linspace(0, g.Ly, g.Ny, endpoint=False, dtype=np.float64)
    g.X, g.Y = np.meshgrid(g.x, g.y)
    g.kx = 2 * np.pi * np.fft.rfftfreq(g.Nx, d=g.dx).astype(np.float64)
    g.ky = 2 * np.pi * np.fft.fftfreq(g.Ny, d=g.dy).astype(np.float64)
    g.KX, g.KY = np.meshgrid(g.kx, g.ky)
    g.K2 = g.KX**2 + g.KY**2
    return g


class _DerivativeKernels:
    pass


def BuildDerivativeKernels(geometry):
    """d_dx = i KX, d_dy = i KY, d2_dx2, d2_dy2 and the spectral Laplacian -(KX^2 + KY^2)."""
    k = _DerivativeKernels()
    k.Geometry = geometry
    k.d_dx = 1j * geometry.KX
    k.d_dy = 1j * geometry.KY
    k.d2_dx2 = -geometry.KX**2
    k.d2_dy2 = -geometry.KY**2
    k.d2_dlap = -(geometry.KX**2 + geometry.KY**2)
    return k


def kernel_local(geometry, w):
    """Gaussian smoothing symbol exp(-k^2 w^2 / 2)."""
    return np.exp(-((geometry.KX)**2 + (geometry.KY)**2) * w**2 / 2)


def kernel_etd(kLinear, dt, *, small_z_tol=2e-5):
    """Host FP64 ETD1 coefficients (E, P) with the small-|z| Taylor guard; the
    device tables use the same formulas (built in the library)."""
    z = kLinear * dt
    e = np.exp(z)
    with np.errstate(divide="ignore", invalid="ignore"):
        p = np.where(np.abs(z) > small_z_tol, (e - 1) / kLinear, dt * (1 + z / 2 + z**2 / 6))
    return e, p
