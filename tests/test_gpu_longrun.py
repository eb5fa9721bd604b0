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
