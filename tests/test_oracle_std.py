"""The FP64 oracle against the reference's recorded values and closed forms (standard PFC).

Every assertion here comes from the reference's own tests
(models/stdpfc_cpu/test_regression.py, tests/test_smoke_models.py) or from
mathematics that a plausible slip (sign of the Laplacian, a dropped 1/2, a
transposed k-grid, a wrong normalisation) would break.
"""
import math

import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

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

# models/stdpfc_cpu/test_regression.py:64-66
REFERENCE_FREE_ENERGY = -0.11237092135766821
REFERENCE_STEP = 200

MODE_NX = 64
MODE_L = 28 * A_LAT


def _sim(nx=132, ny=None, lx=14*A_LAT, ly=8*H_LAT, beta=BETA, eps=EPSILON, g=G, dt=DT, scheme="etd1"):
    ny = nx if ny is None else ny
    geom = O.build_geometry(nx, ny, lx, ly)
    p = O.Params(model="std", Gamma=1.0, beta=beta, q0=Q0, epsilon=eps, g=g, dt=dt, scheme=scheme)
    return O.Simulation(p, geom), geom


def test_recorded_free_energy_step_200():
    """models/stdpfc_cpu/test_regression.py:87-115 (rtol 1e-10, mass < 1e-12)."""
    sim, geom = _sim()
    sim.set_field(F.GeneratePerfectCrystalInitialWithCentralVacancy(geom, MEAN_S, MEAN_V, A0=AMP, q0=Q0))
    m0 = sim.mean()
    sim.step(REFERENCE_STEP)
    assert abs(sim.mean() - m0) < 1e-12
    assert sim.free_energy() == pytest.approx(REFERENCE_FREE_ENERGY, rel=1e-10)


def test_free_energy_single_mode():
    """Closed form for n = nbar + 2a cos(kx) (models/stdpfc_cpu/test_regression.py:153-174)."""
    n_bar, amp, mode = -0.3, 0.07, 4
    sim, geom = _sim(nx=MODE_NX, lx=MODE_L, ly=MODE_L)
    k = 2 * np.pi * mode / MODE_L
    sim.set_field(n_bar + 2 * amp * np.cos(k * geom.X))
    expected = ((BETA / 2) * (n_bar**2 * Q0**4 + 2 * amp**2 * (Q0**2 - k**2)**2)
                + (EPSILON / 2) * (n_bar**2 + 2 * amp**2)
                + (G / 3) * (n_bar**3 + 6 * n_bar * amp**2)
                + 0.25 * (n_bar**4 + 12 * n_bar**2 * amp**2 + 6 * amp**4))
    assert sim.free_energy() == pytest.approx(expected, rel=1e-12, abs=1e-14)


def test_free_energy_single_mode_along_y():
    """Same closed form with the mode along y: pins the ky grid (fftfreq) and the
    Hermitian handling of the kx = 0 column, which the x-mode does not touch."""
    n_bar, amp, mode = 0.1, 0.05, 5
    sim, geom = _sim(nx=MODE_NX, ny=48, lx=MODE_L, ly=0.75*MODE_L)
    k = 2 * np.pi * mode / (0.75*MODE_L)
    sim.set_field(n_bar + 2 * amp * np.cos(k * geom.Y))
    expected = ((BETA / 2) * (n_bar**2 * Q0**4 + 2 * amp**2 * (Q0**2 - k**2)**2)
                + (EPSILON / 2) * (n_bar**2 + 2 * amp**2)
                + (G / 3) * (n_bar**3 + 6 * n_bar * amp**2)
                + 0.25 * (n_bar**4 + 12 * n_bar**2 * amp**2 + 6 * amp**4))
    assert sim.free_energy() == pytest.approx(expected, rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("scheme", ["etd1", "etdrk2"])
@pytest.mark.parametrize("mode", [3, 4, 5])
def test_etd_exact_in_linear_regime(mode, scheme):
    """A 1e-6 mode grows as exp(L t) (models/stdpfc_cpu/test_regression.py:177-199).
    ETDRK2 integrates the linear part exactly too."""
    amp = 1e-6
    sim, geom = _sim(nx=MODE_NX, lx=MODE_L, ly=MODE_L, scheme=scheme)
    k = 2 * np.pi * mode / MODE_L
    sim.set_field(amp * np.cos(k * geom.X))
    lin = -k**2 * (BETA * (Q0**2 - k**2)**2 + EPSILON)
    sim.step(100)
    got = float(np.abs(np.fft.rfft2(sim.field)[0, mode])) / (MODE_NX * MODE_NX / 2)
    assert got == pytest.approx(amp * np.exp(lin * DT * 100), rel=1e-8)


def _removal_offsets(mean_density):
    s = G / 3
    energy = (BETA * Q0**4 * s * (mean_density + s / 2)
              + (EPSILON * s - 2 * s**3) * mean_density
              + 0.5 * EPSILON * s**2 - 1.25 * s**4)
    potential = EPSILON * s - 2 * s**3 + BETA * Q0**4 * s
    return s, energy, potential


def test_cubic_removal_exact_at_fixed_field():
    """models/stdpfc_cpu/test_regression.py:235-270: shifting n by g/3 moves f and mu by constants."""
    sim_a, geom = _sim(nx=MODE_NX, lx=MODE_L, ly=MODE_L)
    rng = np.random.default_rng(7)
    field = MEAN_S + 0.3 * np.fft.irfft2(
        np.fft.rfft2(rng.normal(size=geom.X.shape)) * np.exp(-(geom.KX**2 + geom.KY**2)), s=geom.X.shape)
    s, e_off, mu_off = _removal_offsets(float(field.mean()))
    sim_b, _ = _sim(nx=MODE_NX, lx=MODE_L, ly=MODE_L, eps=EPSILON - G**2/3, g=0.0)
    sim_a.set_field(field)
    sim_b.set_field(field + s)
    assert sim_b.free_energy() - sim_a.free_energy() == pytest.approx(e_off, rel=1e-10, abs=1e-15)
    dmu = sim_b.chemical_potential()[0] - sim_a.chemical_potential()[0]
    assert np.ptp(dmu) < 1e-11
    assert float(dmu.mean()) == pytest.approx(mu_off, rel=1e-9)


def test_cubic_removal_converges_first_order():
    """models/stdpfc_cpu/test_regression.py:273-300."""
    s = G / 3
    gaps = []
    for dt in (0.1, 0.05, 0.025):
        sim_a, geom = _sim(dt=dt)
        init = F.GeneratePerfectCrystalInitialConditions(geom, MEAN_S, A0=AMP, q0=Q0)
        sim_b, _ = _sim(dt=dt, eps=EPSILON - G**2/3, g=0.0)
        sim_a.set_field(init)
        sim_b.set_field(init + s)
        nst = int(round(2.0 / dt))
        sim_a.step(nst)
        sim_b.step(nst)
        gaps.append(float(np.abs(sim_b.field - (sim_a.field + s)).max()))
    ratios = [gaps[i] / gaps[i + 1] for i in range(2)]
    assert all(r > 1.8 for r in ratios), (gaps, ratios)
    assert gaps[-1] < 1e-4


def test_etdrk2_second_order():
    """ETDRK2 error against a fine-dt run falls ~4x per dt halving; ETD1's ~2x."""
    def run(dt, scheme):
        sim, geom = _sim(dt=dt, scheme=scheme)
        sim.set_field(F.GeneratePerfectCrystalInitialWithCentralVacancy(geom, MEAN_S, MEAN_V, A0=AMP, q0=Q0))
        sim.step(int(round(1.6 / dt)))
        return sim.field
    ref = run(0.00625, "etdrk2")
    e2 = [np.abs(run(dt, "etdrk2") - ref).max() for dt in (0.1, 0.05, 0.025)]
    e1 = [np.abs(run(dt, "etd1") - ref).max() for dt in (0.1, 0.05)]
    r2 = [e2[0] / e2[1], e2[1] / e2[2]]
    assert all(3.5 < r < 4.6 for r in r2), (e2, r2)
    assert 1.6 < e1[0] / e1[1] < 2.4, e1


def test_relaxes_monotonically_264():
    """tests/test_smoke_models.py:261-295 (stdpfc: 100 steps on 264^2, FE strictly
    decreasing, mass drift < 1e-12)."""
    lx, ly = 7 * 4 * A_LAT, 4 * 4 * H_LAT
    beta, eps, g = C12 * 0.5, C12 * (-0.9 + 2*(-2.0) + 4), C34 * (-2.0 + 2.5)
    sim, geom = _sim(nx=264, lx=lx, ly=ly, beta=beta, eps=eps, g=g)
    sim.set_field(F.GeneratePerfectCrystalInitialWithCentralVacancy(
        geom, C14 * (0.55 - 1), C14 * (0.30 - 1), A0=C14 * 0.15, q0=Q0))
    fe, mass = [sim.free_energy()], [sim.mean()]
    for _ in range(100):
        sim.step()
        fe.append(sim.free_energy())
        mass.append(sim.mean())
    assert (np.diff(fe) < 0).all()
    assert max(mass) - min(mass) < 1e-12


@pytest.mark.parametrize("mx,my", [(3, 0), (0, 2), (4, 5)])
def test_local_field_single_mode_closed_form(mx, my):
    """Gaussian smoothing of one Fourier mode: local_field(cos(kx x) cos(ky y))
    = exp(-k^2 w^2 / 2) cos(kx x) cos(ky y), k^2 = kx^2 + ky^2, w = 2 pi
    (kernels.py:49-52 applied as in Get_local_n, stdpfc_cpu/simulation.py:250-261).
    A w-for-w^2 or missing-half slip changes the factor by orders of magnitude
    at these k (exp(-k^2 w^2/2) ~ 1e-3 .. 1e-30)."""
    Lx, Ly = 64.0, 48.0
    geom = O.build_geometry(64, 48, Lx, Ly)
    kx, ky = 2 * np.pi * mx / Lx, 2 * np.pi * my / Ly
    f = np.cos(kx * geom.X) * np.cos(ky * geom.Y)
    sim = O.Simulation(O.Params(model="std", epsilon=-0.3, g=0.0, dt=DT), geom)
    sim.set_field(f)
    w = 2 * np.pi
    expect = math.exp(-(kx**2 + ky**2) * w**2 / 2) * f
    got = sim.local_field()
    assert np.abs(got - expect).max() <= 1e-15 + 1e-12 * np.abs(expect).max()
    assert np.abs(expect).max() > 1e-35  # the factor itself is resolved, not zero


def test_geometry_contract():
    """tests/test_smoke_models.py:298-319."""
    geom = O.build_geometry(64, 64, 32.0, 32.0)
    assert np.allclose(geom.K2, geom.KX**2 + geom.KY**2)
    assert np.allclose(O.laplacian(geom), -(geom.KX**2 + geom.KY**2))
    assert np.isclose(O.kernel_local(geom, w=2 * np.pi).max(), 1.0)
    e, p = O.kernel_etd(O.laplacian(geom), DT)
    assert np.isfinite(e).all() and np.isfinite(p).all()
    assert np.isclose(p.flat[0], DT)
    # k-grid layout: ky in fftfreq order on axis 0, kx in rfftfreq order on axis 1
    assert geom.KX.shape == (64, 33)
    assert geom.ky[-1] < 0 and geom.kx[-1] > 0


def test_etdrk2_coefficient_against_high_precision():
    """P2 = dt (e^z - 1 - z)/z^2 against a 50-digit Decimal evaluation."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    zs = np.array([-30.0, -1.5, -0.1000001, -0.0999999, -1e-3, -2e-5, -1e-9, 0.0, 1e-7, 0.05, 0.5])
    dt = 0.5
    got = O.kernel_etdrk2(zs / dt, dt)
    for z, g in zip(zs, got):
        if z == 0.0:
            want = dt * 0.5
        else:
            Z = Decimal(float(z))
            want = float(Decimal(dt) * ((Z.exp() - 1 - Z) / (Z * Z)))
        assert g == pytest.approx(want, rel=5e-15), z


def test_exact_mean_and_peaks():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(16, 12)) * 1e8
    a[0, 0] = 1e16
    a[0, 1] = -1e16
    assert O.exact_mean(a) == pytest.approx(float(np.sum(a[1:]) + np.sum(a[0, 2:])) / a.size, rel=1e-9)
    # brute force: a field built from two Fourier modes peaks on those bins
    geom = O.build_geometry(16, 12, 16.0, 12.0)
    f = 0.3 * np.cos(2 * np.pi * 3 * geom.X / 16.0) + 0.2 * np.cos(2 * np.pi * 2 * geom.Y / 12.0)
    assert O.structure_factor_peaks(f, k=3) == sorted([(0, 3), (2, 0), (10, 0)])


@pytest.mark.parametrize("scheme", ["etd1", "etdrk2"])
def test_external_potential_forcing_is_exact_in_the_linear_regime(scheme):
    # d_t n_hat = L n_hat + lap V_hat with N negligible (g = 0, |n| ~ 1e-9):
    # n_hat(t) = e^{Lt} n_hat0 + (e^{Lt} - 1)/L lap V_hat, which both schemes
    # reproduce exactly for constant forcing; the k = 0 mode is untouched
    geom = O.build_geometry(32, 24, 32 * np.pi / 4, 24 * np.pi / 4)
    # Gamma != 1, so a Gamma wrongly applied to the potential term would show
    p = O.Params(model="std", Gamma=0.7, beta=1.0, q0=1.0, epsilon=0.2, g=0.0, dt=0.3, scheme=scheme)
    sim = O.Simulation(p, geom)
    rng = np.random.default_rng(8)
    n0 = 1e-9 * rng.standard_normal((24, 32))
    V = 1e-9 * (np.sin(2 * np.pi * 2 * geom.X / geom.Lx) + 0.5 * np.cos(2 * np.pi * 3 * geom.Y / geom.Ly)
                + 0.2 * rng.standard_normal((24, 32)))
    sim.set_field(n0)
    sim.set_external_potential(V)
    K = 40
    sim.step(K)
    lap = O.laplacian(geom)
    L = p.Gamma * lap * O.linear_kernel_mu(p.beta, p.q0, p.epsilon, lap)
    t = K * p.dt
    safeL = np.where(L == 0, 1.0, L)
    gain = np.where(L == 0, 0.0, np.expm1(L * t) / safeL)
    exact = np.exp(L * t) * np.fft.rfft2(n0) + gain * lap * np.fft.rfft2(V)
    got = np.fft.rfft2(sim.field)
    assert np.abs(got - exact).max() <= 1e-12 * np.abs(exact).max()
    assert sim.mean() == pytest.approx(n0.mean(), abs=1e-24)
    # long runs settle on the stationary state -lap V_hat / L
    sim.step(4000)
    stat = np.where(L == 0, np.fft.rfft2(n0), -lap * np.fft.rfft2(V) / safeL)
    assert np.abs(np.fft.rfft2(sim.field) - stat).max() <= 1e-10 * np.abs(stat).max()
