"""FP64 CPU oracle for the pseudo-spectral ETD step of the standard and log PFC models.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the `pfc_hip` package, the
HIP library) imports this module; only `tests/`, `__graft_entry__.smoke()` and
the `cpu_baseline` leg of `bench.py` use it, and only as the checker or as the
timed CPU baseline.

It is a NumPy/SciPy restatement of the reference's FP64 CPU path
(`models/stdpfc_cpu/simulation.py`, `models/logpfc_cpu/simulation.py`) with
`numpy.fft` / `scipy.fft` (pocketfft) in place of pyFFTW, which is absent here.
Every function names the reference lines it follows.  The step keeps the
reference's order and arithmetic:

    Batch1  N = nonlin(n);  [n_hat, N_hat] = rfft2([n, N])
    Batch2  mu_hat = L_mu n_hat + N_hat; local_mu_hat = G mu_hat (diagnostic only)
    Batch3  n_next_hat = E n_hat + P (Gamma lap N_hat);  n = irfft2(n_next_hat)

Batch2 does not feed the update; the oracle computes it only on request
(`chemical_potential`).  Extensions the reference does not have -- ETDRK2, the
logPFC log clamp and M-shift as options, a k=0 pin -- are documented where
they appear, with what pins them.

How each function is pinned (see tests/test_oracle_*.py):
  geometry, derivative_kernels, kernel_etd, kernel_local  -- the reference's
      geometry-contract assertions (tests/test_smoke_models.py:298-319) and the
      closed forms below;
  StdPFC step + free_energy -- the recorded regression value -0.11237092135766821
      at step 200 on 132^2 (models/stdpfc_cpu/test_regression.py:64-115), the
      single-mode free energy, ETD1 linear exactness, cubic-removal identities and
      first-order convergence (same file :153-300);
  LogPFC step + free_energy -- the recorded values at steps 10/100/1000 on 28x24
      (models/logpfc_cpu/test_regression.py:28-66) and the liquid closed form;
  kernel_etdrk2 + ETDRK2 step -- linear-regime exactness and a measured
      second-order convergence ratio (~4) under dt halving;
  clamp / M-shift / pin -- invariants (mass, positivity) and the identity
      that an M shift changes an ETD1 trajectory only at O(dt);
  external potential (logpfc_extv's lap V_hat forcing) -- the closed-form
      solution of the linear problem with constant forcing, which ETD1 and
      ETDRK2 reproduce exactly, and the stationary state -lap V_hat / L.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

try:  # scipy.fft gives a `workers=` knob for the multi-core CPU baseline
    import scipy.fft as _sfft
except Exception:  # pragma: no cover - scipy is present in this image
    _sfft = None


# ---------------------------------------------------------------------------
# geometry and spectral operators
# ---------------------------------------------------------------------------
@dataclass
class Geometry:
    """2D periodic spectral geometry (discretization/spectral/geometry_2d.py:38-88)."""
    Nx: int
    Ny: int
    Lx: float
    Ly: float
    dx: float = 0.0
    dy: float = 0.0
    x: np.ndarray = field(default=None, repr=False)
    y: np.ndarray = field(default=None, repr=False)
    X: np.ndarray = field(default=None, repr=False)
    Y: np.ndarray = field(default=None, repr=False)
    kx: np.ndarray = field(default=None, repr=False)
    ky: np.ndarray = field(default=None, repr=False)
    KX: np.ndarray = field(default=None, repr=False)
    KY: np.ndarray = field(default=None, repr=False)
    K2: np.ndarray = field(default=None, repr=False)


def build_geometry(Nx, Ny, Lx, Ly):
    """Restates Build2DGeometry (geometry_2d.py:70-88): kx from rfftfreq, ky from
    fftfreq, both times 2 pi; meshgrids in C order (Ny, Nx) / (Ny, Nx//2+1)."""
    g = Geometry(Nx=int(Nx), Ny=int(Ny), Lx=Lx, Ly=Ly)
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


def laplacian(geom):
    """Spectral Laplacian symbol d2_dlap = -(KX^2 + KY^2) (kernels.py:45)."""
    return -(geom.KX**2 + geom.KY**2)


def kernel_local(geom, w):
    """Gaussian smoothing symbol exp(-k^2 w^2 / 2) (kernels.py:49-52)."""
    return np.exp(-((geom.KX)**2 + (geom.KY)**2) * w**2 / 2)


def kernel_etd(k_linear, dt, small_z_tol=2e-5):
    """ETD1 coefficients (kernels.py:55-71): E = exp(L dt);
    P = (E - 1)/L where |L dt| > tol, else dt (1 + z/2 + z^2/6)."""
    z = k_linear * dt
    e = np.exp(z)
    with np.errstate(divide='ignore', invalid='ignore'):
        p = np.where(np.abs(z) > small_z_tol, (e - 1) / k_linear,
                     dt * (1 + z / 2 + z**2 / 6))
    return e, p


def kernel_etdrk2(k_linear, dt):
    """Second ETDRK2 coefficient P2 = (e^z - 1 - z) / (L^2 dt), z = L dt.

    Not in the reference (it has ETD1 only, model.toml `order = 1`); this is the
    ETD2RK scheme of Cox & Matthews, J. Comput. Phys. 176 (2002) 430, eq. (22):
        a       = E u + P F(u)
        u_next  = a + P2 (F(a) - F(u))
    P2 = dt * phi2(z) with phi2(z) = (e^z - 1 - z)/z^2.  For |z| < 0.1 phi2 is
    summed from its Taylor series (16 terms, truncation < 1e-30); elsewhere from
    expm1, whose absolute error eps*|z| becomes eps/|z| < 3e-15 relative.
    Pinned by linear-regime exactness and the second-order convergence ratio.
    """
    z = np.asarray(k_linear * dt, dtype=np.float64)
    small = np.abs(z) < 0.1
    series = np.zeros_like(z)
    term = np.full_like(z, 0.5)          # z^0 / 2!
    for j in range(16):
        series = series + term
        term = term * z / (j + 3)
    with np.errstate(divide='ignore', invalid='ignore'):
        direct = (np.expm1(z) - z) / (z * z)
    phi2 = np.where(small, series, direct)
    return dt * phi2


def linear_kernel_mu(beta, q0, epsilon, lap, M=0.0):
    """L_mu = beta (q0^2 + lap)^2 + epsilon (+ M) (stdpfc_cpu/simulation.py:95-102;
    M shift as logpfc_extv/simulation.py:100-107)."""
    return beta * (q0**2 + lap)**2 + epsilon + M


def linear_kernel_n(Gamma, beta, q0, epsilon, lap, M=0.0):
    """L = Gamma lap L_mu (stdpfc_cpu/simulation.py:104-108)."""
    return Gamma * lap * linear_kernel_mu(beta, q0, epsilon, lap, M)


# ---------------------------------------------------------------------------
# FFT provider (numpy.fft by default; scipy.fft with worker threads on request)
# ---------------------------------------------------------------------------
class _FFT:
    def __init__(self, lib="numpy", workers=1):
        self.lib = lib
        self.workers = workers
        if lib == "scipy" and _sfft is None:
            raise RuntimeError("scipy.fft unavailable")

    def rfft2(self, a):
        if self.lib == "scipy":
            return _sfft.rfft2(a, workers=self.workers)
        return np.fft.rfft2(a)

    def irfft2(self, a, s):
        if self.lib == "scipy":
            return _sfft.irfft2(a, s=s, workers=self.workers)
        return np.fft.irfft2(a, s=s)


# ---------------------------------------------------------------------------
# model parameters
# ---------------------------------------------------------------------------
@dataclass
class Params:
    """Model parameters (stdpfc_cpu/simulation.py:72-92; logpfc_cpu/simulation.py:60-84).

    model      'std' (field n) or 'log' (field phi)
    scheme     'etd1' (the reference) or 'etdrk2' (new; Cox & Matthews 2002)
    clamp_log  logPFC only: False = the reference's unclamped ln(phi)
               (logpfc_cpu/simulation.py:321-331); True = the documented clamp
               where(phi > minPhi, ln phi, ln minPhi) (logpfc_extv/simulation.py:537-539)
    M          logPFC stabilising shift: +M on L_mu and -M phi on N
               (logpfc_extv/simulation.py:100-116, :538)
    pin_mean   None: no pin (the reference's stdpfc/logpfc); a float: the k=0
               coefficient of every update is set to N*pin_mean
               (logpfc_extv/simulation.py:577 pins it to the previous value)
    """
    model: str = "std"
    Gamma: float = 1.0
    beta: float = 1.0
    q0: float = 1.0
    epsilon: float = 1.0
    g: float = 1.0
    dt: float = 0.01
    minPhi: float = 1e-6
    smallZTol: float = 2e-5
    scheme: str = "etd1"
    clamp_log: bool = False
    M: float = 0.0
    pin_mean: float | None = None


# ---------------------------------------------------------------------------
# pointwise physics
# ---------------------------------------------------------------------------
def nonlinear_mu(field_r, p: Params):
    """Real-space nonlinear chemical potential.

    std: g n^2 + n^3                      (stdpfc_cpu/simulation.py:323-333)
    log: ln(phi) + g phi^2 + phi^3 [- M phi]   (logpfc_cpu/simulation.py:321-331;
         clamp and M shift as logpfc_extv/simulation.py:536-539)
    """
    f = field_r
    if p.model == "std":
        return p.g * (f**2) + (f**3)
    if p.clamp_log:
        with np.errstate(divide="ignore", invalid="ignore"):
            lg = np.where(f > p.minPhi, np.log(f), np.log(p.minPhi))
    else:
        with np.errstate(divide="ignore", invalid="ignore"):
            lg = np.log(f)
    out = lg + p.g * (f**2) + (f**3)
    if p.M != 0.0:
        out = out - p.M * f
    return out


def local_free_energy(field_r, p: Params):
    """Local (non-gradient) free-energy density.

    std: eps/2 n^2 + g/3 n^3 + n^4/4                 (stdpfc_cpu/simulation.py:281)
    log: phi ln phi + eps/2 phi^2 + g/3 phi^3 + phi^4/4 (logpfc_cpu/simulation.py:279)
    """
    f = field_r
    poly = (p.epsilon / 2.0) * f**2 + (p.g / 3.0) * f**3 + (1 / 4.0) * f**4
    if p.model == "std":
        return poly
    with np.errstate(divide="ignore", invalid="ignore"):
        return f * np.log(f) + poly


# ---------------------------------------------------------------------------
# the simulation
# ---------------------------------------------------------------------------
class Simulation:
    """One FP64 instance stepped exactly as the reference's CPU path.

    The field lives in real space (`self.field`, shape (Ny, Nx)).  `step()` is
    TimeStep (stdpfc_cpu/simulation.py:238-248) without Batch2 (which does not
    feed the update); `free_energy()` is CalculateFreeEnergy (:263-285) and
    returns the box mean of f, as the reference does.
    """

    def __init__(self, p: Params, geom: Geometry, fft="numpy", workers=1):
        self.p = p
        self.geom = geom
        self.fft = _FFT(fft, workers)
        self.t = 0.0
        self.lap = laplacian(geom)
        self.L_mu = linear_kernel_mu(p.beta, p.q0, p.epsilon, self.lap, p.M)
        self.L = p.Gamma * self.lap * self.L_mu
        self.E, self.P = kernel_etd(self.L, p.dt, small_z_tol=p.smallZTol)
        self.P2 = kernel_etdrk2(self.L, p.dt) if p.scheme == "etdrk2" else None
        self.G = kernel_local(geom, w=2 * np.pi)
        self.field = np.zeros((geom.Ny, geom.Nx), dtype=np.float64)
        self.lapV_hat = None

    def set_external_potential(self, V):
        """External potential V (or None): the forcing gains lap V_hat, without
        Gamma (logpfc_extv/simulation.py:557-577); constant, so in ETDRK2 it
        cancels from the second stage."""
        self.lapV_hat = None if V is None else self.lap * self.fft.rfft2(np.asarray(V, dtype=np.float64))

    # -- state -------------------------------------------------------------
    def set_field(self, f):
        self.field = np.array(f, dtype=np.float64, copy=True).reshape(self.geom.Ny, self.geom.Nx)

    def mean(self):
        return float(self.field.mean())

    # -- one nonlinear evaluation ------------------------------------------
    def _forcing_hat(self, field_r):
        """Gamma lap rfft2(N(field)) -- the ETD forcing (stdpfc_cpu/simulation.py:365-381)."""
        N_hat = self.fft.rfft2(nonlinear_mu(field_r, self.p))
        return self.p.Gamma * self.lap * N_hat

    def _pin(self, u_hat):
        if self.p.pin_mean is not None:
            u_hat.flat[0] = self.p.pin_mean * self.geom.Nx * self.geom.Ny
        return u_hat

    def step(self, nsteps=1):
        s = (self.geom.Ny, self.geom.Nx)
        for _ in range(int(nsteps)):
            u = self.field
            u_hat = self.fft.rfft2(u)
            F_u = self._forcing_hat(u)
            # ETD1 (stdpfc_cpu/simulation.py:365-397)
            a_hat = self.E * u_hat + self.P * (F_u if self.lapV_hat is None else F_u + self.lapV_hat)
            if self.p.scheme == "etd1":
                self.field = self.fft.irfft2(self._pin(a_hat), s)
            elif self.p.scheme == "etdrk2":
                a_hat = self._pin(a_hat)
                a = self.fft.irfft2(a_hat, s)
                F_a = self._forcing_hat(a)
                u_next = a_hat + self.P2 * (F_a - F_u)
                self.field = self.fft.irfft2(self._pin(u_next), s)
            else:
                raise ValueError(f"unknown scheme {self.p.scheme!r}")
            self.t += self.p.dt
        return self

    # -- reductions ---------------------------------------------------------
    def free_energy(self):
        """CalculateFreeEnergy (stdpfc_cpu/simulation.py:263-285; log :259-283)."""
        p = self.p
        n = self.field
        n_hat = self.fft.rfft2(n)
        lin_f = (p.beta / 2.0) * (p.q0**2 + self.lap)**2
        lin = self.fft.irfft2(lin_f * n_hat, n.shape)
        f = n * lin + local_free_energy(n, p)
        return float(f.mean())

    def chemical_potential(self):
        """Batch2 of the reference (stdpfc_cpu/simulation.py:349-363): mu and <mu>."""
        n = self.field
        n_hat = self.fft.rfft2(n)
        N_hat = self.fft.rfft2(nonlinear_mu(n, self.p))
        mu_hat = self.L_mu * n_hat + N_hat
        s = n.shape
        return self.fft.irfft2(mu_hat, s), self.fft.irfft2(self.G * mu_hat, s)

    def local_field(self):
        """Get_local_n (stdpfc_cpu/simulation.py:250-261): Gaussian-smoothed field."""
        return self.fft.irfft2(self.G * self.fft.rfft2(self.field), self.field.shape)


# ---------------------------------------------------------------------------
# logPFC stabilising shift (logpfc_extv/simulation.py:332-373)
# ---------------------------------------------------------------------------
def optimal_shift(phi_min, phi_max, g, minPhi):
    """M = (N(phi_max) - N(phi_min)) / (phi_max - phi_min), N = ln phi + g phi^2 + phi^3,
    with phi_min replaced by minPhi when the field is not positive (extv :336-344)."""
    lo = phi_min if phi_min > 0 else minPhi
    n_hi = g * phi_max**2 + phi_max**3 + math.log(phi_max)
    n_lo = g * lo**2 + lo**3 + math.log(lo)
    return (n_hi - n_lo) / (phi_max - lo)


# ---------------------------------------------------------------------------
# exact reductions and structure factor
# ---------------------------------------------------------------------------
def exact_mean(a):
    """Correctly rounded mean via math.fsum (brute-force reduction reference)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    return math.fsum(a.tolist()) / a.size


def structure_factor_peaks(field_r, k=6, exclude_dc=True):
    """Indices (ky, kx) of the k largest |rfft2|^2 entries, sorted by power then index.

    Used for the structure-factor parity check: the same crystal must put its
    Bragg peaks on the same half-spectrum bins.
    """
    S = np.abs(np.fft.rfft2(np.asarray(field_r, dtype=np.float64)))**2
    if exclude_dc:
        S[0, 0] = -1.0
    flat = np.argsort(-S, axis=None, kind="stable")[:k]
    iy, ix = np.unravel_index(flat, S.shape)
    return sorted(zip(iy.tolist(), ix.tolist()))
