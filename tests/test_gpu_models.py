"""The reference's model tests, run against the hip drop-in modules (needs a GPU).

Each test follows a reference test line for line where one exists (cited);
only the module and builder names change (stdpfc_cpu -> stdpfc_hip,
BuildNumpyCPU_Simulation -> BuildHipGPU_Simulation).
"""
import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

pytestmark = pytest.mark.gpu

Q0 = 1.0
A_LAT = 4 * np.pi / 3**0.5 / Q0
H_LAT = 4 * np.pi / Q0
C12, C34, C14 = (3/4)**0.5, (3/4)**0.75, (3/4)**0.25
BETA = C12 * 0.5
EPSILON = C12 * (-0.9 + 2*(-2.0) + 4)
G = C34 * (-2.0 + 2.5)
MEAN_S = C14 * (0.55 - 1)
MEAN_V = C14 * (0.30 - 1)
AMP = C14 * 0.19
DT = 0.1
MODE_NX = 64
MODE_L = 28 * A_LAT


def _std():
    from pfc_hip.models.stdpfc_hip import simulation as sim_mod
    return sim_mod


def _build(beta=BETA, epsilon=EPSILON, g=G, mean=MEAN_S, dt=DT, nx=132, lx=14 * A_LAT, ly=8 * H_LAT, **kw):
    sim_mod = _std()
    geometry = sim_mod.Build2DGeometry(Nx=nx, Ny=nx, Lx=lx, Ly=ly)
    params = sim_mod.BuildStdPFCModelParameters(Gamma=1.0, beta=beta, q0=Q0, epsilon=epsilon, g=g, dt=dt,
                                                meanDensity=mean)
    return sim_mod.BuildHipGPU_Simulation(params, geometry, **kw), geometry


def test_stdpfc_recorded_free_energy(hip_device):
    """models/stdpfc_cpu/test_regression.py:87-115."""
    sim, geometry = _build()
    sim.State.n[...] = F.GeneratePerfectCrystalInitialWithCentralVacancy(geometry, MEAN_S, MEAN_V, A0=AMP, q0=Q0)
    mass_start = float(sim.State.n.mean())
    for _ in range(200):
        sim.TimeStep()
    assert abs(float(sim.State.n.mean()) - mass_start) < 1e-12
    assert float(sim.CalculateFreeEnergy()) == pytest.approx(-0.11237092135766821, rel=1e-10)
    assert sim.t == pytest.approx(20.0)


def test_stdpfc_run_equals_timestep_loop(hip_device):
    a, geometry = _build()
    b, _ = _build()
    init = F.GeneratePerfectCrystalInitialWithCentralVacancy(geometry, MEAN_S, MEAN_V, A0=AMP, q0=Q0)
    a.State.n[...] = init
    b.State.n[...] = init
    for _ in range(25):
        a.TimeStep()
    b.run(25)
    assert np.array_equal(np.asarray(a.State.n), np.asarray(b.State.n))
    assert a.t == b.t


def test_free_energy_single_mode(hip_device):
    """models/stdpfc_cpu/test_regression.py:153-174 (rel 1e-12)."""
    n_bar, amp, mode = -0.3, 0.07, 4
    sim, geometry = _build(nx=MODE_NX, lx=MODE_L, ly=MODE_L, mean=0.0)
    k = 2 * np.pi * mode / MODE_L
    sim.State.n[...] = n_bar + 2 * amp * np.cos(k * geometry.X)
    expected = ((BETA / 2) * (n_bar**2 * Q0**4 + 2 * amp**2 * (Q0**2 - k**2)**2)
                + (EPSILON / 2) * (n_bar**2 + 2 * amp**2)
                + (G / 3) * (n_bar**3 + 6 * n_bar * amp**2)
                + 0.25 * (n_bar**4 + 12 * n_bar**2 * amp**2 + 6 * amp**4))
    assert float(sim.CalculateFreeEnergy()) == pytest.approx(expected, rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("scheme", ["etd1", "etdrk2"])
@pytest.mark.parametrize("mode", [3, 4, 5])
def test_etd_is_exact_in_the_linear_regime(hip_device, mode, scheme):
    """models/stdpfc_cpu/test_regression.py:177-199 (rel 1e-8)."""
    amp = 1e-6
    sim, geometry = _build(nx=MODE_NX, lx=MODE_L, ly=MODE_L, mean=0.0, scheme=scheme)
    k = 2 * np.pi * mode / MODE_L
    sim.State.n[...] = amp * np.cos(k * geometry.X)
    k_lap = -k**2
    linear = k_lap * (BETA * (Q0**2 + k_lap)**2 + EPSILON)
    for _ in range(100):
        sim.TimeStep()
    spectrum = np.fft.rfft2(sim.State.n)
    got = float(np.abs(spectrum[0, mode])) / (MODE_NX * MODE_NX / 2)
    assert got == pytest.approx(amp * np.exp(linear * DT * 100), rel=1e-8)


def test_cubic_removal_is_exact_at_fixed_field(hip_device):
    """models/stdpfc_cpu/test_regression.py:235-270, including _Batch1/_Batch2 -> State.mu."""
    sim_a, geometry = _build(nx=MODE_NX, lx=MODE_L, ly=MODE_L, mean=0.0)
    rng = np.random.default_rng(7)
    field = MEAN_S + 0.3 * np.fft.irfft2(
        np.fft.rfft2(rng.normal(size=geometry.X.shape)) * np.exp(-(geometry.KX**2 + geometry.KY**2)),
        s=geometry.X.shape)
    mean_density = float(field.mean())
    s = G / 3
    energy = (BETA * Q0**4 * s * (mean_density + s / 2) + (EPSILON * s - 2 * s**3) * mean_density
              + 0.5 * EPSILON * s**2 - 1.25 * s**4)
    potential = EPSILON * s - 2 * s**3 + BETA * Q0**4 * s
    sim_b, _ = _build(nx=MODE_NX, lx=MODE_L, ly=MODE_L, epsilon=EPSILON - G**2/3, g=0.0, mean=mean_density + s)
    sim_a.State.n[...] = field
    sim_b.State.n[...] = field + s
    delta_f = float(sim_b.CalculateFreeEnergy()) - float(sim_a.CalculateFreeEnergy())
    assert delta_f == pytest.approx(energy, rel=1e-10, abs=1e-15)
    for sim in (sim_a, sim_b):
        sim._Batch1_CalcNonlinMu()
        sim._Batch2_CalcMu_Terms()
    delta_mu = sim_b.State.mu - sim_a.State.mu
    assert np.ptp(delta_mu) < 1e-11
    assert float(delta_mu.mean()) == pytest.approx(potential, rel=1e-9)


def test_chemical_potential_and_local_field_match_oracle(hip_device):
    sim, geometry = _build(nx=MODE_NX, lx=MODE_L, ly=MODE_L)
    rng = np.random.default_rng(2)
    f = MEAN_S + 0.1 * rng.normal(size=geometry.X.shape)
    sim.State.n[...] = f
    sim._Batch2_CalcMu_Terms()
    ref = O.Simulation(O.Params(model="std", beta=BETA, epsilon=EPSILON, g=G, dt=DT),
                       O.build_geometry(MODE_NX, MODE_NX, MODE_L, MODE_L))
    ref.set_field(f)
    mu, lmu = ref.chemical_potential()
    assert np.abs(sim.State.mu - mu).max() < 1e-12
    assert np.abs(sim.State.local_mu - lmu).max() < 1e-12
    assert np.abs(sim.Get_local_n() - ref.local_field()).max() < 1e-13


def test_logpfc_recorded_free_energies(hip_device):
    """models/logpfc_cpu/test_regression.py:40-66."""
    from pfc_hip.models.logpfc_hip import simulation as sim_mod
    Q, MEAN = 0.9648639799649456, 0.6326
    a, h = 4 * np.pi / (np.sqrt(3) * Q), 4 * np.pi / Q
    geometry = sim_mod.Build2DGeometry(Nx=28, Ny=24, Lx=2 * a, Ly=h)
    params = sim_mod.BuildLogPFCModelParameters(Gamma=1.0, beta=0.5, q0=1.0, epsilon=-0.9, g=-2.0, dt=0.1,
                                                meanDensity=MEAN)
    sim = sim_mod.BuildHipGPU_Simulation(params, geometry)
    X, Y = geometry.X, geometry.Y
    phi = MEAN * (1 + 0.15 * (2 * np.cos(np.sqrt(3) / 2 * Q * X) * np.cos(Q * Y / 2) + np.cos(Q * Y)))
    phi += 1e-3 * np.random.default_rng(7).standard_normal(phi.shape)
    phi += MEAN - phi.mean()
    sim.State.phi[...] = phi
    ref = {10: -0.5162767800170736, 100: -0.554845641895177, 1000: -0.5548456430520675}
    done = 0
    for step in sorted(ref):
        for _ in range(step - done):
            sim.TimeStep()
        done = step
        assert sim.CalculateFreeEnergy() == pytest.approx(ref[step], rel=1e-10), step
    assert sim.phi.mean() == pytest.approx(MEAN, abs=1e-12)


def _march(sim, field, nsteps=100):
    fe = [float(sim.CalculateFreeEnergy())]
    mass = [float(getattr(sim.State, field).mean())]
    for _ in range(nsteps):
        sim.TimeStep()
        fe.append(float(sim.CalculateFreeEnergy()))
        mass.append(float(getattr(sim.State, field).mean()))
    return np.array(fe), np.array(mass)


def test_stdpfc_relaxes(hip_device):
    """tests/test_smoke_models.py:261-277 (264^2, 100 steps, FE strictly decreasing, mass < 1e-12)."""
    sim, geometry = _build(nx=264, lx=28 * A_LAT, ly=16 * H_LAT, beta=C12 * 0.5,
                           epsilon=C12 * (-0.9 + 2*(-2.0) + 4), g=C34 * (-2.0 + 2.5))
    sim.State.n[...] = F.GeneratePerfectCrystalInitialWithCentralVacancy(
        geometry, C14 * (0.55 - 1), C14 * (0.30 - 1), A0=C14 * 0.15, q0=Q0)
    fe, mass = _march(sim, "n")
    assert np.isfinite(fe).all() and np.isfinite(mass).all()
    assert mass.max() - mass.min() < 1e-12
    assert (np.diff(fe) < 0).all() and fe[-1] < fe[0]


@pytest.mark.parametrize("stabilise", [False, True])
def test_logpfc_relaxes(hip_device, stabilise):
    """tests/test_smoke_models.py:135-171: 264^2, 100 steps; finite, phi > 0, mass < 1e-12;
    FE strictly decreasing for the plain model (the stabilised one, like logpfc_extv,
    is exempt from monotonicity: its operator changes during the run, :30-45)."""
    from pfc_hip.models.logpfc_hip import simulation as sim_mod
    geometry = sim_mod.Build2DGeometry(Nx=264, Ny=264, Lx=28 * A_LAT, Ly=16 * H_LAT)
    params = sim_mod.BuildLogPFCModelParameters(Gamma=1.0, beta=0.5, q0=Q0, epsilon=-0.9, g=-2.0, dt=DT,
                                                meanDensity=0.55)
    sim = sim_mod.BuildHipGPU_Simulation(params, geometry, stabilise=stabilise)
    sim.State.phi[...] = F.GeneratePerfectCrystalInitialWithCentralVacancy(geometry, 0.55, 0.3, A0=0.15, q0=Q0)
    fe, mass = _march(sim, "phi")
    assert np.isfinite(fe).all() and np.isfinite(mass).all()
    assert float(sim.State.phi.min()) > 0.0
    assert mass.max() - mass.min() < 1e-12
    if not stabilise:
        assert (np.diff(fe) < 0).all()
    else:
        assert sim.State.M != 0.0


def test_logpfc_stabilised_matches_oracle_shift(hip_device):
    """With M fixed, the device M-shifted split equals the oracle's to 1e-11 over 50 steps."""
    from pfc_hip.plan import Plan
    geom = O.build_geometry(96, 64, 96 * np.pi / 4, 64 * np.pi / 4)
    phi0 = 0.8 + 0.05 * np.random.default_rng(9).uniform(-1, 1, size=(64, 96))
    M = O.optimal_shift(phi0.min(), phi0.max(), 1.0, 1e-6)
    sim = O.Simulation(O.Params(model="log", dt=0.05, clamp_log=True, M=M), geom)
    sim.set_field(phi0)
    sim.step(50)
    pl = Plan(model="log", Nx=96, Ny=64, Lx=geom.Lx, Ly=geom.Ly, dt=0.05, clamp_log=True, M=M)
    pl.set_field(phi0).step(50)
    assert np.abs(pl.get_field().cpu().numpy() - sim.field).max() < 1e-11


def test_hip_backend_contract(hip_device):
    import torch
    from pfc_hip.backends import hip_rocm as B
    assert B.name == "hip" and B.available()
    x = np.random.default_rng(4).normal(size=(2, 48, 64))
    X = B.rfftn(torch.as_tensor(x, device="cuda"))
    assert np.allclose(B.to_numpy(X), np.fft.rfft2(x), rtol=0, atol=1e-12)
    back = B.to_numpy(B.irfftn(X, s=(48, 64)))
    assert np.abs(back - x).max() < 1e-14
