"""The FP64 oracle against the reference's recorded values and closed forms (log PFC)."""
import math

import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

# models/logpfc_cpu/test_regression.py:28-37
REFERENCE_FREE_ENERGY = {10: -0.5162767800170736, 100: -0.554845641895177, 1000: -0.5548456430520675}
Q = 0.9648639799649456
MEAN = 0.6326


def _regression_sim(**kw):
    a, h = 4 * np.pi / (np.sqrt(3) * Q), 4 * np.pi / Q
    geom = O.build_geometry(28, 24, 2 * a, h)
    p = O.Params(model="log", Gamma=1.0, beta=0.5, q0=1.0, epsilon=-0.9, g=-2.0, dt=0.1, **kw)
    sim = O.Simulation(p, geom)
    X, Y = geom.X, geom.Y
    phi = MEAN * (1 + 0.15 * (2 * np.cos(np.sqrt(3) / 2 * Q * X) * np.cos(Q * Y / 2) + np.cos(Q * Y)))
    phi += 1e-3 * np.random.default_rng(7).standard_normal(phi.shape)
    phi += MEAN - phi.mean()
    sim.set_field(phi)
    return sim


def test_recorded_free_energy_steps_10_100_1000():
    """models/logpfc_cpu/test_regression.py:58-66 (rtol 1e-10)."""
    sim = _regression_sim()
    done = 0
    for step in sorted(REFERENCE_FREE_ENERGY):
        sim.step(step - done)
        done = step
        assert sim.free_energy() == pytest.approx(REFERENCE_FREE_ENERGY[step], rel=1e-10), step
    assert sim.mean() == pytest.approx(MEAN, abs=1e-12)


def test_clamp_is_inert_above_min_phi():
    """With phi > minPhi everywhere the clamped log equals ln phi bit for bit."""
    a = _regression_sim()
    b = _regression_sim(clamp_log=True)
    a.step(50)
    b.step(50)
    assert np.array_equal(a.field, b.field)


def test_clamp_guards_nonpositive_phi():
    p = O.Params(model="log", g=-2.0, clamp_log=True, minPhi=1e-6)
    phi = np.array([-0.5, 0.0, 1e-7, 0.3])
    got = O.nonlinear_mu(phi, p)
    want = np.where(phi > 1e-6, np.log(np.where(phi > 0, phi, 1.0)), math.log(1e-6)) - 2.0 * phi**2 + phi**3
    assert np.array_equal(got, want)
    assert np.isfinite(got).all()


def test_liquid_free_energy_closed_form():
    """Uniform phi: f = phi ln phi + (beta q0^4 + eps)/2 phi^2 + g/3 phi^3 + phi^4/4
    (models/logpfc_cpu/phase_diagram.py:322-333)."""
    geom = O.build_geometry(32, 24, 20.0, 15.0)
    beta, q0, eps, g, phi = 0.7, 1.1, -0.4, 0.3, 0.8
    sim = O.Simulation(O.Params(model="log", beta=beta, q0=q0, epsilon=eps, g=g), geom)
    sim.set_field(np.full((24, 32), phi))
    want = phi * math.log(phi) + 0.5 * (beta * q0**4 + eps) * phi**2 + g / 3 * phi**3 + phi**4 / 4
    assert sim.free_energy() == pytest.approx(want, rel=1e-14)


def test_m_shift_changes_trajectory_only_at_first_order():
    """+M on L_mu and -M phi on N is the same continuous dynamics; under ETD1 the
    split moves, so trajectories differ at O(dt) (cf. the cubic-removal pair,
    models/stdpfc_cpu/test_regression.py:205-222)."""
    gaps = []
    for dt in (0.1, 0.05, 0.025):
        a = _regression_sim()
        b = _regression_sim(M=1.7)
        a.p.dt = b.p.dt = dt
        a.__init__(a.p, a.geom)
        b.__init__(b.p, b.geom)
        src = _regression_sim().field
        a.set_field(src)
        b.set_field(src)
        n = int(round(2.0 / dt))
        a.step(n)
        b.step(n)
        gaps.append(float(np.abs(a.field - b.field).max()))
        assert abs(b.mean() - MEAN) < 1e-12
    r = [gaps[0] / gaps[1], gaps[1] / gaps[2]]
    assert all(1.7 < x < 2.3 for x in r), (gaps, r)


def test_pin_holds_mean_exactly():
    sim = _regression_sim(pin_mean=MEAN)
    sim.step(20)
    assert abs(sim.mean() - MEAN) < 1e-15


@pytest.mark.parametrize("lo,hi,g", [(0.2, 1.4, -2.0), (0.5, 0.6, 1.0), (0.05, 2.5, 0.5)])
def test_optimal_shift_levels_the_shifted_nonlinearity(lo, hi, g):
    """What the shift is for (logpfc_extv/simulation.py:332-373): with M =
    optimal_shift(lo, hi), the shifted nonlinearity N(phi) - M phi, evaluated
    by the step's own nonlinear_mu, takes equal values at the window's ends;
    and M lies between N' at the two ends (mean value theorem, N convex or
    not, the secant slope is attained inside)."""
    M = O.optimal_shift(lo, hi, g, 1e-6)
    p = O.Params(model="log", g=g, clamp_log=True, M=M)
    v = O.nonlinear_mu(np.array([lo, hi]), p)
    assert abs(v[1] - v[0]) <= 1e-13 * max(1.0, abs(v[0]))
    xs = np.linspace(lo, hi, 20001)
    dN = 1 / xs + 2 * g * xs + 3 * xs**2
    assert dN.min() <= M <= dN.max()


def test_optimal_shift_floors_a_non_positive_minimum():
    """phi_min <= 0 is replaced by minPhi (extv :336-344): the shifted
    nonlinearity with the log clamped at minPhi is level between minPhi and
    phi_max."""
    minPhi, hi, g = 1e-6, 1.4, -2.0
    M = O.optimal_shift(-0.1, hi, g, minPhi)
    assert M == O.optimal_shift(minPhi, hi, g, minPhi)
    p = O.Params(model="log", g=g, clamp_log=True, M=M, minPhi=minPhi)
    v = O.nonlinear_mu(np.array([minPhi, hi, -0.1]), p)
    assert abs(v[1] - v[0]) <= 1e-12 * abs(v[0])


def test_log_relaxes_monotonically_264():
    """tests/test_smoke_models.py:135-151 (logpfc_cpu, 100 steps, FE strictly decreasing)."""
    a_lat, h_lat = 4 * np.pi / 3**0.5, 4 * np.pi
    geom = O.build_geometry(264, 264, 28 * a_lat, 16 * h_lat)
    sim = O.Simulation(O.Params(model="log", beta=0.5, epsilon=-0.9, g=-2.0, dt=0.1), geom)
    sim.set_field(F.GeneratePerfectCrystalInitialWithCentralVacancy(geom, 0.55, 0.3, A0=0.15, q0=1.0))
    fe, mass = [sim.free_energy()], [sim.mean()]
    for _ in range(100):
        sim.step()
        fe.append(sim.free_energy())
        mass.append(sim.mean())
    assert (np.diff(fe) < 0).all()
    assert max(mass) - min(mass) < 1e-12
    assert sim.field.min() > 0
