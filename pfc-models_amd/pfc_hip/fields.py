"""Initial conditions on the host (FP64 NumPy), uploaded as-is to the device.

The first two generators restate the reference's deterministic crystal ICs
(fields/initial_conditions_2d.py:80-144), which its recorded regression values
are pinned on.  The rest build the synthetic inputs of the benchmark
workloads; each draws from `numpy.random.default_rng(seed)` (PCG64) so the
device and the CPU checker start from bit-identical fields.
"""
from __future__ import annotations

import math

import numpy as np

A_LATTICE_Q1 = 4 * np.pi / math.sqrt(3)   # triangular lattice spacing at q0 = 1
H_LATTICE_Q1 = 4 * np.pi                  # rectangular-cell row spacing at q0 = 1


def _one_mode_vectors(q0):
    q1 = q0 * np.array([-(3**0.5) / 2, -0.5])
    q2 = q0 * np.array([3**0.5 / 2, -0.5])
    q3 = q0 * np.array([0, 1])
    return q1, q2, q3


def _triangular(X, Y, mean, A0, q0):
    q1, q2, q3 = _one_mode_vectors(q0)
    return (
        mean
        + A0 * np.exp(1.0j * (q1[0] * X + q1[1] * Y))
        + A0 * np.exp(-1.0j * (q1[0] * X + q1[1] * Y))
        + A0 * np.exp(1.0j * (q2[0] * X + q2[1] * Y))
        + A0 * np.exp(-1.0j * (q2[0] * X + q2[1] * Y))
        + A0 * np.exp(1.0j * (q3[0] * X + q3[1] * Y))
        + A0 * np.exp(-1.0j * (q3[0] * X + q3[1] * Y))
    ).real


def GeneratePerfectCrystalInitialConditions(geometry, meanDensity, A0, q0=1.0):
    """One-mode triangular crystal (fields/initial_conditions_2d.py:80-104)."""
    return _triangular(geometry.X, geometry.Y, meanDensity, A0, q0).astype(np.float64)


def GeneratePerfectCrystalInitialWithCentralVacancy(geometry, meanDensity_S, meanDensity_V, A0, q0=1.0):
    """Crystal with one disk of radius a/2 set to meanDensity_V at the central
    lattice site (fields/initial_conditions_2d.py:107-144)."""
    X, Y = geometry.X, geometry.Y
    a = 4 * np.pi / 3**0.5 / q0
    h = 4 * np.pi / q0
    x_center = (geometry.Lx / a // 2) * a
    y_center = (geometry.Ly / h // 2) * h
    phi = _triangular(X, Y, meanDensity_S, A0, q0)
    phi[((X - x_center) ** 2 + (Y - y_center) ** 2) < (a**2 / 4)] = meanDensity_V
    return phi.astype(np.float64)


# ---------------------------------------------------------------------------
# benchmark workloads
# ---------------------------------------------------------------------------
def uniform_noise(shape, mean, half_width, seed):
    """mean + U(-half_width, half_width) i.i.d. per site."""
    rng = np.random.default_rng(seed)
    return mean + rng.uniform(-half_width, half_width, size=shape)


def one_mode_amplitude(r, psibar):
    """A = (4/5)(psibar + sqrt(-15 r - 36 psibar^2)/3): the one-mode triangular
    amplitude of the standard PFC model at q = sqrt(3)/2 scaling."""
    return 0.8 * (psibar + math.sqrt(-15.0 * r - 36.0 * psibar**2) / 3.0)


def seeded_crystals(geometry, r, psibar, n_seeds=8, radius_pts=20, noise_std=0.01, seed=2):
    """Liquid psibar + N(0, noise_std^2) holding `n_seeds` circular crystal seeds.

    Draw order from default_rng(seed): the (Ny, Nx) noise, then seed centres
    U[0,Lx) x U[0,Ly) as an (n_seeds, 2) array, then rotation angles U[0, pi/3).
    Inside a seed (periodic minimum-image distance < radius_pts * dx) psi is the
    one-mode profile psibar + A (cos(q x') cos(q y'/sqrt3) - 0.5 cos(2 q y'/sqrt3))
    with q = sqrt(3)/2 and (x', y') the seed-frame coordinates.
    """
    rng = np.random.default_rng(seed)
    Ny, Nx = geometry.Ny, geometry.Nx
    psi = psibar + rng.normal(0.0, noise_std, size=(Ny, Nx))
    centres = rng.uniform(0.0, 1.0, size=(n_seeds, 2)) * np.array([geometry.Lx, geometry.Ly])
    angles = rng.uniform(0.0, np.pi / 3, size=n_seeds)
    A = one_mode_amplitude(r, psibar)
    q = math.sqrt(3) / 2
    R = radius_pts * geometry.dx
    X, Y = geometry.X, geometry.Y
    for (cx, cy), th in zip(centres, angles):
        ddx = (X - cx + geometry.Lx / 2) % geometry.Lx - geometry.Lx / 2
        ddy = (Y - cy + geometry.Ly / 2) % geometry.Ly - geometry.Ly / 2
        inside = ddx**2 + ddy**2 < R**2
        c, s = math.cos(th), math.sin(th)
        xp = c * ddx + s * ddy
        yp = -s * ddx + c * ddy
        prof = psibar + A * (np.cos(q * xp) * np.cos(q * yp / math.sqrt(3))
                             - 0.5 * np.cos(2 * q * yp / math.sqrt(3)))
        psi[inside] = prof[inside]
    return psi


def log_noise(shape, psibar=-0.2, half_width=0.02, seed=3, floor=1e-6):
    """logPFC density phi = 1 + psi with psi = psibar + U(-w, w), clipped to
    psi > -1 + floor (so phi > floor)."""
    psi = uniform_noise(shape, psibar, half_width, seed)
    psi = np.maximum(psi, -1.0 + floor)
    return 1.0 + psi


def commensurate_box(Nx, Ny, dx=np.pi / 4, q0=1.0):
    """Lx, Ly holding an integer number of triangular periods: Lx = round(Nx dx / a) a,
    Ly = round(Ny dx / h) h (the rectangular cell of models/stdpfc_cpu/phase_diagram.py:161-167)."""
    a = 4 * np.pi / (math.sqrt(3) * q0)
    h = 4 * np.pi / q0
    mx = max(1, int(round(Nx * dx / a)))
    my = max(1, int(round(Ny * dx / h)))
    return mx * a, my * h
