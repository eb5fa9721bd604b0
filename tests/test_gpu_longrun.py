"""Long runs against the FP64 oracle (needs a GPU).

FP64: relative free energy per site <= 1e-9 at every checkpoint and the same
structure-factor peak positions at the end.  The round-off floor this is
measured against: perturbing the initial field of the 128^2 run by one ulp
moves its free energy by 3e-15 relative over 2000 steps (the dynamics relax
without amplifying round-off), so 1e-9 is a real bound.  FP32 bulk (FP64 ETD
tables, combine and reductions): relative free energy <= 1e-6.
"""
import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

pytestmark = pytest.mark.gpu


def _compare(geom, p, psi0, steps, every, precision=64, fe_tol=1e-9):
    from pfc_hip.plan import Plan
    sim = O.Simulation(p, geom)
    sim.set_field(psi0)
    pl = Plan(Nx=geom.Nx, Ny=geom.Ny, Lx=geom.Lx, Ly=geom.Ly, epsilon=p.epsilon, g=p.g, dt=p.dt,
              scheme=p.scheme, precision=precision)
    pl.set_field(psi0)
    m0 = pl.mean()
    worst = 0.0
    for _ in range(steps // every):
        sim.step(every)
        pl.step(every)
        fe, mean, _, _ = pl.reduce()
        ref = sim.free_energy()
        worst = max(worst, abs(fe - ref) / abs(ref))
        assert abs(mean - m0) <= (1e-12 if precision == 64 else 1e-7)
    assert worst <= fe_tol, worst
    got = pl.get_field().cpu().numpy()
    assert O.structure_factor_peaks(got, 12) == O.structure_factor_peaks(sim.field, 12)
    return worst


def test_std_128_2000_steps(hip_device):
    """128^2, r = psibar = -0.25, ETD1 dt = 0.5, 2000 steps, FE every 100 steps."""
    geom = O.build_geometry(128, 128, 32 * np.pi, 32 * np.pi)
    p = O.Params(model="std", epsilon=-0.25, g=0.0, dt=0.5)
    _compare(geom, p, F.uniform_noise((128, 128), -0.25, 0.05, seed=1), 2000, 100)


def test_seeded_crystals_etdrk2_256(hip_device):
    """The 1024^2 benchmark's initial condition at 256^2 (8 one-mode seeds in a
    psibar = -0.3 liquid, r = -0.3), ETDRK2 dt = 0.5, 1000 steps."""
    n = 256
    geom = O.build_geometry(n, n, n * np.pi / 4, n * np.pi / 4)
    p = O.Params(model="std", epsilon=-0.3, g=0.0, dt=0.5, scheme="etdrk2")
    _compare(geom, p, F.seeded_crystals(geom, -0.3, -0.3), 1000, 100)


def test_std_128_fp32_bulk(hip_device):
    """FP32 bulk on the 128^2 r = psibar = -0.25 run: relative FE <= 1e-6 against
    the FP64 oracle at every checkpoint, the mean held by the k = 0 pin, and
    the same structure-factor peaks."""
    geom = O.build_geometry(128, 128, 32 * np.pi, 32 * np.pi)
    p = O.Params(model="std", epsilon=-0.25, g=0.0, dt=0.5)
    _compare(geom, p, F.uniform_noise((128, 128), -0.25, 0.05, seed=1), 2000, 100, precision=32, fe_tol=1e-6)


def test_logpfc_2048_fp32_stabilised_10000_steps(hip_device):
    """The logPFC 2048^2 workload: builder defaults (Gamma = beta = q0 = eps =
    g = 1, dt = 0.01, minPhi = 1e-6, models/logpfc/simulation.py:51-75), dx =
    pi/4, psibar = -0.2 (phi = 0.8) + U(-0.02, 0.02) clipped to psi > -1 +
    1e-6, ETD1 with the stabilised split (M re-fitted when the phi range
    drifts by > 10 %), FP32 bulk, 10 000 steps.
    * The first 20 steps follow the FP64 oracle with the same shift to
      max|dphi| <= 2e-6 and relative free energy <= 1e-6.
    * After 10 000 steps: finite, phi > 0, the mean constant (k = 0 pin); the
      FP64 device reduction of the final field (through an FP64 plan) equals
      the oracle's free energy of that field to 1e-12; the free energy fell
      and lies above the uniform liquid's closed form at the realised mean
      (the linearly stable state this phi relaxes to,
      models/logpfc_cpu/phase_diagram.py:324-333) by less than 1e-5 relative."""
    import os
    from pfc_hip.models.logpfc_hip import simulation as sim_mod
    from pfc_hip.models.logpfc_hip import phase_diagram as LAD
    from pfc_hip.plan import Plan
    n = 2048
    L = n * np.pi / 4
    psi0 = np.clip(F.uniform_noise((n, n), -0.2, 0.02, seed=3), -1 + 1e-6, None)
    phi0 = 1.0 + psi0
    mean0 = O.exact_mean(phi0)
    geometry = sim_mod.Build2DGeometry(Nx=n, Ny=n, Lx=L, Ly=L)
    params = sim_mod.BuildLogPFCModelParameters(meanDensity=mean0)
    sim = sim_mod.BuildHipGPU_Simulation(params, geometry, precision=32, stabilise=True)
    sim.State.phi[...] = phi0
    fe0 = float(sim.CalculateFreeEnergy())
    sim.run(20)
    # fitted once from the FP32 field's min and max (not re-fitted: the range shrinks)
    M = sim.State.M
    assert abs(M - O.optimal_shift(phi0.min(), phi0.max(), 1.0, 1e-6)) <= 1e-6 * abs(M)
    geom = O.build_geometry(n, n, L, L)
    ref = O.Simulation(O.Params(model="log", dt=0.01, clamp_log=True, M=M, pin_mean=mean0), geom, fft="scipy",
                       workers=min(16, os.cpu_count() or 1))
    ref.set_field(phi0)
    ref.step(20)
    got = sim.State.phi.get()
    mean20 = sim._plan.reduce()[1]
    assert np.abs(got - ref.field).max() <= 2e-6
    assert abs(float(sim.CalculateFreeEnergy()) - ref.free_energy()) <= 1e-6 * abs(ref.free_energy())
    sim.run(10000 - 20)
    phi = sim.State.phi.get()
    assert np.isfinite(phi).all() and phi.min() > 0.0
    fe, mean, _, _ = sim._plan.reduce()
    # the k = 0 coefficient is pinned every step to the FP64 mean; the field
    # holds it up to the FP32 rounding of the stored rows (re-pinned every
    # step, so it does not drift): well inside one FP32 ulp of the mean
    ulp = float(np.spacing(np.float32(mean0)))
    assert abs(mean - mean0) <= 0.05 * ulp and abs(mean - mean20) <= 0.05 * ulp, (mean, mean20, mean0)
    pl = Plan(model="log", Nx=n, Ny=n, Lx=L, Ly=L, dt=0.01, clamp_log=True).set_field(phi)
    fe64 = pl.free_energy()
    chk = O.Simulation(O.Params(model="log", dt=0.01, clamp_log=True), geom, fft="scipy",
                       workers=min(16, os.cpu_count() or 1))
    chk.set_field(phi)
    assert abs(fe64 - chk.free_energy()) <= 1e-12 * abs(chk.free_energy())
    liq = LAD.liquid_free_energy(mean, {"beta": 1.0, "g": 1.0, "q0": 1.0, "v0": 1.0}, 1.0)
    assert fe < fe0
    assert 0.0 <= fe - liq <= 1e-5 * abs(liq), (fe, liq)
