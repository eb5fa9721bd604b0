"""The external-potential forcing of logpfc_extv (P lap V_hat in the update,
models/logpfc_extv/simulation.py:557-577) on the HIP path against the oracle
(needs a GPU): FP64 max|d field| <= 1e-10 after 100 steps on the resident
stepper (28 x 24), the generic column kernel (128^2, ETD1 and ETDRK2), the
linear closed form, batch slots, removal, and the logpfc_extv_hip model."""
import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

pytestmark = pytest.mark.gpu


def _potential(geom, amp):
    return amp * (np.tanh((np.abs(geom.X - geom.Lx / 2) - geom.Lx / 6) / 2.0)
                  + 0.3 * np.sin(2 * np.pi * geom.Y / geom.Ly))


def _run(model, Nx, Ny, scheme, steps, V, x0, eps, g, beta, dt, M=0.0):
    from pfc_hip.plan import Plan
    a = 4 * np.pi / 3 ** 0.5
    # the rectangular cell of the triangular lattice where the grid is a whole
    # number of cells (28 x 24), else dx = dy = pi/4
    Lx = (Nx // 14) * a if Nx % 14 == 0 else Nx * np.pi / 4
    Ly = (Ny // 24) * 4 * np.pi if Ny % 24 == 0 else Ny * np.pi / 4
    geom = O.build_geometry(Nx, Ny, Lx, Ly)
    Vf = _potential(geom, V)
    p = O.Params(model=model, Gamma=0.8, beta=beta, q0=1.0, epsilon=eps, g=g, dt=dt, scheme=scheme, M=M,
                 clamp_log=(model == "log"), pin_mean=float(np.mean(x0)))
    ref = O.Simulation(p, geom)
    ref.set_field(x0)
    ref.set_external_potential(Vf)
    ref.step(steps)
    pl = Plan(model=model, Nx=Nx, Ny=Ny, Lx=Lx, Ly=Ly, Gamma=0.8, beta=beta, q0=1.0, epsilon=eps, g=g, dt=dt,
              M=M, precision=64, scheme=scheme, clamp_log=(model == "log")).set_field(x0)
    pl.set_external_potential(Vf)
    pl.step(steps)
    return pl, ref, geom, Vf


@pytest.mark.parametrize("case", [
    dict(model="log", Nx=28, Ny=24, scheme="etd1", eps=-0.9, g=-2.0, beta=0.5, dt=0.05, mean=0.63, V=0.05),
    dict(model="std", Nx=128, Ny=128, scheme="etd1", eps=-0.25, g=0.0, beta=1.0, dt=0.5, mean=-0.25, V=0.02),
    dict(model="std", Nx=128, Ny=128, scheme="etdrk2", eps=-0.25, g=0.0, beta=1.0, dt=0.5, mean=-0.25, V=0.02),
], ids=["log28x24-resident", "std128-etd1", "std128-etdrk2"])
def test_potential_step_matches_the_oracle(hip_device, case):
    x0 = F.uniform_noise((case["Ny"], case["Nx"]), case["mean"], 0.03, seed=21)
    pl, ref, _, _ = _run(case["model"], case["Nx"], case["Ny"], case["scheme"], 100, case["V"], x0, case["eps"],
                         case["g"], case["beta"], case["dt"])
    got = pl.get_field().cpu().numpy()
    d = np.abs(got - ref.field).max()
    print(f"\n{case['model']} {case['Nx']}x{case['Ny']} {case['scheme']}: max |d| {d:.3e} "
          f"(field moved {np.abs(ref.field - x0).max():.2e})", flush=True)
    assert d <= 1e-10
    assert pl.mean() == pytest.approx(float(np.mean(x0)), abs=1e-14)
    assert pl.free_energy() == pytest.approx(ref.free_energy(), rel=1e-12)


def test_potential_linear_closed_form(hip_device):
    # std, g = 0, |n| ~ 1e-9: the step is linear with the constant source
    # lap V_hat, whose exact solution ETD1 reproduces
    from pfc_hip.plan import Plan
    Nx, Ny, dt, K = 32, 24, 0.3, 40
    geom = O.build_geometry(Nx, Ny, Nx * np.pi / 4, Ny * np.pi / 4)
    rng = np.random.default_rng(8)
    n0 = 1e-9 * rng.standard_normal((Ny, Nx))
    V = 1e-9 * (np.sin(2 * np.pi * 2 * geom.X / geom.Lx) + 0.2 * rng.standard_normal((Ny, Nx)))
    pl = Plan(model="std", Nx=Nx, Ny=Ny, Lx=geom.Lx, Ly=geom.Ly, Gamma=0.7, beta=1.0, q0=1.0, epsilon=0.2, g=0.0,
              dt=dt, precision=64).set_field(n0)
    pl.set_external_potential(V).step(K)
    lap = O.laplacian(geom)
    L = 0.7 * lap * O.linear_kernel_mu(1.0, 1.0, 0.2, lap)
    safeL = np.where(L == 0, 1.0, L)
    gain = np.where(L == 0, 0.0, np.expm1(L * K * dt) / safeL)
    exact = np.exp(L * K * dt) * np.fft.rfft2(n0) + gain * lap * np.fft.rfft2(V)
    got = np.fft.rfft2(pl.get_field().cpu().numpy())
    assert np.abs(got - exact).max() <= 1e-12 * np.abs(exact).max()


def test_potential_per_slot_and_removal(hip_device):
    from pfc_hip.plan import Batch, Plan
    Nx, Ny = 28, 24
    a = 4 * np.pi / 3 ** 0.5
    kw = dict(model="log", Nx=Nx, Ny=Ny, Lx=2 * a, Ly=4 * np.pi, Gamma=1.0, beta=0.5, q0=1.0, epsilon=-0.9, g=-2.0,
              dt=0.05, precision=64)
    geom = O.build_geometry(Nx, Ny, kw["Lx"], kw["Ly"])
    V = _potential(geom, 0.05)
    x0 = F.uniform_noise((Ny, Nx), 0.63, 0.03, seed=22)
    b = Batch(3, **kw)
    for s in range(3):
        b.set_field(s, x0)
    b.set_external_potential(1, V)
    b.set_external_potential(2, V)
    b.set_external_potential(2, None)   # set and removed
    b.step(200)
    plain = Plan(**kw).set_field(x0).step(200).get_field().cpu().numpy()
    withV = Plan(**kw).set_field(x0).set_external_potential(V).step(200).get_field().cpu().numpy()
    assert np.array_equal(b.get_field(0).cpu().numpy(), plain)
    assert np.array_equal(b.get_field(2).cpu().numpy(), plain)
    assert np.array_equal(b.get_field(1).cpu().numpy(), withV)
    assert np.abs(withV - plain).max() > 1e-4


def test_logpfc_extv_model(hip_device):
    # the drop-in module: clamped log, M-shifted split, potential API
    from pfc_hip.models.logpfc_extv_hip import simulation as X
    geom = X.BuildGeometry(28, 24, 2 * 4 * np.pi / 3 ** 0.5, 4 * np.pi)
    m = X.BuildLogPFCModel_withVacancyForce_Parameters(beta=0.5, epsilon=-0.9, g=-2.0, dt=0.05, V0=0.05, potWidth=2.0)
    assert m.smallZTol == 1e-8 and m.V0 == 0.05
    sim = X.BuildHipGPU_Simulation(m, geom)
    phi0 = F.uniform_noise((24, 28), 0.63, 0.03, seed=23)
    sim.State.phi[...] = phi0
    sim._externalV.SetPotential_alongX(0.05, 2.0, geom.Lx / 3, 2 * geom.Lx / 3)
    for _ in range(50):
        sim.TimeStep()
    phi = np.asarray(sim.phi.get())
    assert abs(phi.mean() - phi0.mean()) < 1e-13 and phi.min() > 0 and sim.State.M != 0
    # a smooth potential for the force: -d/dx against a central difference
    sim._externalV.SetGaussianDipolePotential(0.05, (geom.Lx / 3, geom.Ly / 2), (2 * geom.Lx / 3, geom.Ly / 2), 1.5)
    F_ = sim.GetForce()
    assert F_.shape == (2, 24, 28)
    Vx = sim._externalV.V
    fd = -(np.roll(Vx, -1, axis=1) - np.roll(Vx, 1, axis=1)) / (2 * geom.dx)
    assert np.abs(F_[0] - fd).max() < 0.05 * np.abs(fd).max()
    sim._CalcLapV()
    assert sim._externalV.lap_V.shape == (24, 28)
    sim._externalV.clear()
    sim.TimeStep()
