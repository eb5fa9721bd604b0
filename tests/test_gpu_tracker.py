"""Tracker reductions: free energy, mean, min and max recorded on the device
every k steps and synchronised once (needs a GPU).

The records must equal, bit for bit, stepping in blocks of k with a reduction
(and a host synchronisation) after each block, the form the reference's
callers use (tests/test_smoke_models.py:104-113); on the 128^2 stdPFC
workload (r = psibar = -0.25, ETD1 dt = 0.5, free energy and mean every 100
of 2000 steps) they follow the FP64 oracle to 1e-9 relative free energy and
the mean stays at its initial value to 1e-12.
"""
import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

pytestmark = pytest.mark.gpu

N = 128
L = N * np.pi / 4


def _plan(precision=64):
    from pfc_hip.plan import Plan
    return Plan(Nx=N, Ny=N, Lx=L, Ly=L, epsilon=-0.25, g=0.0, dt=0.5, precision=precision, scheme="etd1")


def test_plan_records_equal_blockwise_reductions_and_oracle(hip_device):
    psi0 = F.uniform_noise((N, N), -0.25, 0.05, seed=1)
    a = _plan().set_field(psi0)
    rec = a.run_tracked(2000, 100)
    assert rec.shape == (20, 4)
    b = _plan().set_field(psi0)
    blk = []
    for _ in range(20):
        b.step(100)
        blk.append(b.reduce())
    assert np.array_equal(rec, np.array(blk))
    assert np.array_equal(a.get_field().cpu().numpy(), b.get_field().cpu().numpy())
    sim = O.Simulation(O.Params(model="std", epsilon=-0.25, g=0.0, dt=0.5), O.build_geometry(N, N, L, L))
    sim.set_field(psi0)
    m0 = O.exact_mean(psi0)
    for r in range(20):
        sim.step(100)
        ref = sim.free_energy()
        assert abs(rec[r, 0] - ref) <= 1e-9 * abs(ref), r
        assert abs(rec[r, 1] - m0) <= 1e-12
    assert np.all(np.diff(rec[:, 0]) <= 0)  # the free energy never rises


def test_trailing_steps_and_errors(hip_device):
    from pfc_hip import _lib
    psi0 = F.uniform_noise((N, N), -0.25, 0.05, seed=2)
    a = _plan().set_field(psi0)
    rec = a.run_tracked(250, 100)  # two records, then 50 unrecorded steps
    assert rec.shape == (2, 4)
    b = _plan().set_field(psi0).step(250)
    assert np.array_equal(a.get_field().cpu().numpy(), b.get_field().cpu().numpy())
    assert a.run_tracked(40, 100).shape == (0, 4)
    with pytest.raises(_lib.PfcError):
        a.run_tracked(100, 0)


def test_batch_records_per_active_slot(hip_device):
    from pfc_hip.plan import Batch

    def make():
        bt = Batch(3, Nx=N, Ny=N, Lx=L, Ly=L, g=0.0, dt=0.5, precision=32, scheme="etdrk2")
        for s in range(3):
            bt.set_slot(s, -0.35 + 0.1 * s)
            bt.set_field(s, F.uniform_noise((N, N), -0.3, 0.05, seed=10 + s))
        bt.set_active([2, 0])
        return bt

    a, b = make(), make()
    rec = a.run_tracked(300, 50)
    assert rec.shape == (6, 2, 4)
    for r in range(6):
        b.step(50)
        assert np.array_equal(rec[r], b.reduce()), r


def test_simulation_run_records(hip_device):
    from pfc_hip.models.stdpfc_hip import simulation as S
    geom = S.Build2DGeometry(Nx=N, Ny=N, Lx=L, Ly=L)
    params = S.BuildStdPFCModelParameters(Gamma=1.0, beta=1.0, q0=1.0, epsilon=-0.25, g=0.0, dt=0.5,
                                          meanDensity=-0.25)
    sim = S.BuildHipGPU_Simulation(params, geom)
    sim.State.n[...] = F.uniform_noise((N, N), -0.25, 0.05, seed=3)
    rec = sim.run(500, record_every=100)
    assert rec.shape == (5, 4)
    assert sim.t == pytest.approx(250.0)
    assert rec[-1, 0] == pytest.approx(float(sim.CalculateFreeEnergy()), rel=1e-15)
