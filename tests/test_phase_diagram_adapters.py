"""Host side of the phase-diagram adapters (no GPU): parameter schemas, the
uniform-state analytics (pinned on the reference's stored stability-bounds
table, critical temperatures and closed forms) and the 3-mode ansatz and
fitness (pinned by round trips and lattice geometry)."""
import json
import os

import numpy as np
import pytest

from pfc_hip.models import _adapter as A
from pfc_hip.models.logpfc_hip import phase_diagram as LOG
from pfc_hip.models.stdpfc_hip import phase_diagram as STD

HERE = os.path.dirname(os.path.abspath(__file__))
REFS = json.load(open(os.path.join(HERE, "golden", "phase_diagram_refs.json")))
LOG_PARAMS = REFS["log_params"]
CELL = REFS["cell"]
ANSATZ = REFS["ansatz"]


@pytest.mark.parametrize("ad", [LOG, STD], ids=["log", "std"])
def test_validate_params(ad):
    assert ad.validate_params(dict(beta=0.5, g=-2)) == dict(beta=0.5, g=-2.0, v0=1.0, Gamma=1.0, q0=1.0)
    for bad in (dict(beta=0.5, g=-2, v0=0.9), dict(beta=0.5), dict(beta=0.5, g=-2, gamma=1.0)):
        with pytest.raises(ValueError):
            ad.validate_params(bad)
    with pytest.raises(ValueError):
        ad.validate_cell(dict(uNx=14, uNy=24, Mx=2))
    with pytest.raises(ValueError):
        ad.validate_cell(dict(uNx=14, uNy=24, Mx=0, My=1))
    with pytest.raises(ValueError):
        ad.validate_ansatz(dict(mode=1, A1=1, A2=1, A3=1))
    assert ad.validate_cell(CELL) == CELL and ad.validate_ansatz(ANSATZ) == ANSATZ


# -- logPFC analytics --------------------------------------------------------------
def test_log_critical_point():
    c = REFS["log_critical_point"]
    phi_c, eps_c = LOG.critical_point(dict(beta=c["beta"], g=c["g"]))
    assert abs(phi_c - c["phi_c"]) <= c["tol"] and abs(eps_c - c["eps_c"]) <= c["tol"]
    # phi_c minimises h = 1/phi + 2 g phi + 3 phi^2: 6 phi^3 + 2 g phi^2 - 1 = 0 at g = -2
    assert 6 * phi_c ** 3 - 4 * phi_c ** 2 - 1 == pytest.approx(0, abs=1e-14)
    assert eps_c == pytest.approx(-(1 / phi_c - 4 * phi_c + 3 * phi_c ** 2), abs=1e-15)


def test_log_critical_temperature_column():
    # the stored legacy column lies below the exact eps_c by less than 4e-4 at every g
    for beta, g, tab in REFS["critical_temp_table"]:
        _, eps_c = LOG.critical_point(dict(beta=beta, g=g))
        assert 0 < eps_c - tab < 4e-4, (beta, g)


def test_log_spinodal_and_existence_bounds_match_the_stored_table():
    for eps, phi_minus, phi_plus, tri_low, hc_high in REFS["log_stability_bounds"]:
        lo, hi = LOG.spinodal(LOG_PARAMS, eps)
        assert (lo, hi) == pytest.approx((phi_minus, phi_plus), abs=1e-11)
        b = LOG.existence_bounds(LOG_PARAMS, eps)
        assert b["triangle"] == pytest.approx((tri_low, phi_plus), abs=1e-11)
        assert b["honeycomb"] == pytest.approx((phi_minus, hc_high), abs=1e-11)
        assert b["stripe"] == (lo, hi)
        w = LOG.scan_window("triangle", LOG_PARAMS, eps)
        assert w[0] < tri_low and w[1] > phi_plus


def test_log_bands_vanish_at_the_critical_point():
    phi_c, eps_c = LOG.critical_point(LOG_PARAMS)
    assert LOG.spinodal(LOG_PARAMS, eps_c + 1e-9) is None
    assert LOG.existence_bounds(LOG_PARAMS, eps_c + 1e-9) is None
    lo, hi = LOG.spinodal(LOG_PARAMS, eps_c - 1e-6)
    assert lo < phi_c < hi
    # k = 0: the same phi_c, eps shifted by beta q0^4
    beta = LOG_PARAMS["beta"]
    assert LOG.liquid_spinodal(LOG_PARAMS, eps_c - beta + 1e-9) is None
    lo, hi = LOG.liquid_spinodal(LOG_PARAMS, eps_c - beta - 1e-6)
    assert lo < phi_c < hi


@pytest.mark.parametrize("ad, x", [(LOG, 0.7), (STD, -0.3)], ids=["log", "std"])
def test_liquid_chemical_potential_is_the_derivative(ad, x):
    p, eps, h = dict(beta=0.5, g=-2.0), -0.9, 1e-6
    num = (ad.liquid_free_energy(x + h, p, eps) - ad.liquid_free_energy(x - h, p, eps)) / (2 * h)
    assert ad.liquid_chemical_potential(x, p, eps) == pytest.approx(num, rel=1e-8)


# -- stdPFC analytics -----------------------------------------------------------------
def test_std_closed_forms():
    eps = -0.25
    p0 = dict(beta=1.0, g=0.0)
    assert STD.critical_point(dict(beta=1.0, g=-1.5)) == (0.5, 0.75)
    lo, hi = STD.spinodal(p0, eps)
    assert (lo, hi) == pytest.approx((-np.sqrt(-eps / 3), np.sqrt(-eps / 3)), abs=1e-15)
    b = STD.existence_bounds(p0, eps)
    assert b["triangle"][0] == pytest.approx(-np.sqrt(-5 * eps / 12), abs=1e-15)
    assert b["honeycomb"][1] == pytest.approx(np.sqrt(-5 * eps / 12), abs=1e-15)
    p, eps = dict(beta=1.0, g=0.8), -0.4
    for x in STD.spinodal(p, eps):
        assert eps + 2 * 0.8 * x + 3 * x ** 2 == pytest.approx(0, abs=1e-14)
    assert STD.spinodal(p, 0.8 ** 2 / 3 + 1e-9) is None
    for x in STD.liquid_spinodal(p, eps - 1.0):
        assert 1.0 + eps - 1.0 + 2 * 0.8 * x + 3 * x ** 2 == pytest.approx(0, abs=1e-14)
    lo_t, hi_t = STD.existence_bounds(p0, -0.25)["triangle"]
    assert STD.scan_window("triangle", p0, -0.25) == (lo_t - 0.02, hi_t + 0.02)


def test_std_one_mode_saddle_node_is_where_the_hexagon_appears():
    # one-mode hexagons n = nbar + 2 A sum_j cos(k_j . r), |k_j| = q0 = 1, g = 0:
    # <d^2> = 6 A^2, <d^3> = 12 A^3, <d^4> = 90 A^4, so f(A) - f(0) =
    # 3 (eps + 3 nbar^2) A^2 + 12 nbar A^3 + 45/2 A^4.  A nonzero stationary
    # amplitude, 6 (eps + 3 nbar^2) + 36 nbar A + 90 A^2 = 0, exists while the
    # discriminant 36^2 nbar^2 - 4*90*6 (eps + 3 nbar^2) >= 0: it vanishes at
    # the saddle node.  The integrals <d^k> are checked on a grid.
    eps = -0.25
    lo, _ = STD.existence_bounds(dict(beta=1.0, g=0.0), eps)["triangle"]
    assert 1296 * lo ** 2 - 2160 * (eps + 3 * lo ** 2) == pytest.approx(0, abs=1e-12)
    c = dict(uNx=28, uNy=48, Mx=1, My=1)
    d = A.three_mode(STD.STRUCTURES["triangle"], c, 0.0, 1.0, 0.0, 0.0)
    assert [np.mean(d ** k) for k in (2, 3, 4)] == pytest.approx([6.0, 12.0, 90.0], abs=1e-12)


# -- ansatz, fitness and box ------------------------------------------------------------
@pytest.mark.parametrize("name", ["triangle", "honeycomb", "stripe"])
def test_three_mode_fitness_round_trip(name):
    s = STD.STRUCTURES[name]
    c = STD.validate_cell(CELL)
    field = A.three_mode(s, c, -0.2, 0.11 * s.amplitudeSign, 0.023 * s.amplitudeSign, 0.004 * s.amplitudeSign)
    assert field.shape == A.grid_shape(c) == (24, 28)
    assert field.mean() == pytest.approx(-0.2, abs=1e-15)
    fit, a, b, cc = A.fitness(field, s, c)
    assert fit == pytest.approx(1.0, abs=1e-15)
    assert (a, b, cc) == pytest.approx((0.11, 0.023, 0.004), abs=1e-15)
    # relative amplitudes are in units of the mean (the logPFC convention)
    pos = A.three_mode(s, c, 0.8, 0.08 * s.amplitudeSign, 0.01 * s.amplitudeSign, 0.002 * s.amplitudeSign)
    fit, a, b, cc = A.fitness(pos, s, c, relative=True)
    assert fit == pytest.approx(1.0, abs=1e-15) and (a, b, cc) == pytest.approx((0.1, 0.0125, 0.0025), abs=1e-14)


def test_fitness_sees_a_foreign_mode():
    s = STD.STRUCTURES["triangle"]
    c = STD.validate_cell(CELL)
    field = A.three_mode(s, c, 0.0, 0.1, 0.0, 0.0)
    ny, nx = field.shape
    x = np.arange(nx)[None, :] * np.ones((ny, 1))
    extra = 0.01 * np.cos(2 * np.pi * 5 * x / nx)            # not one of the structure's modes
    fit, *_ = A.fitness(field + extra, s, c)
    assert fit == pytest.approx(1 - 0.5 * 0.01 ** 2, abs=1e-15)


@pytest.mark.parametrize("name, expected", [("triangle", 2), ("honeycomb", 4)])
def test_maxima_per_cell(name, expected):
    s = STD.STRUCTURES[name]
    c = dict(uNx=28, uNy=48, Mx=2, My=1)
    field = A.three_mode(s, c, 0.0, 0.1 * s.amplitudeSign, 0.0, 0.0)
    assert A.maxima_per_cell(field, c) == expected == s.maxima_per_cell


def test_box_puts_the_first_shell_at_q00():
    # the triangle's 1-mode set sits on |k| = q00 in the box built for q00
    q00 = 0.93
    c = STD.validate_cell(CELL)
    g = A.box(c, q00)
    (r1, c1), (r2, c2), _ = A.mode_indices(STD.STRUCTURES["triangle"], c)
    kx = 2 * np.pi * c1 / g.Lx
    ky = 2 * np.pi * r1 / g.Ly
    assert np.hypot(kx, ky) == pytest.approx(np.full(4, q00), abs=1e-14)
    kx2, ky2 = 2 * np.pi * c2 / g.Lx, 2 * np.pi * r2 / g.Ly
    assert np.hypot(kx2, ky2) == pytest.approx(np.full(3, np.sqrt(3) * q00), abs=1e-14)


def test_log_ansatz_is_positive_with_the_requested_mean():
    c = STD.validate_cell(CELL)
    for name in ("triangle", "honeycomb", "stripe"):
        big = dict(mode="L2H", A1=3.0, A2=0.0, A3=0.0)   # would go negative unscaled
        f = LOG.initial_field(name, c, big, 0.6)
        assert f.min() == pytest.approx(0.0, abs=1e-15) and f.mean() == pytest.approx(0.6, abs=1e-14)
        small = LOG.initial_field(name, c, ANSATZ, 0.6)
        assert small.min() > 0.59


def _device_rows(field, fit, k2=None):
    """What pfc_diagnose returns, computed with NumPy from the field."""
    spec = np.fft.rfft2(field)
    N = field.size
    full = np.abs(np.fft.fft2(field) / N) ** 2
    m1 = m2 = 0.0
    if k2 is not None:
        m1, m2 = (full * k2).sum(), (full * k2 * k2).sum()
    return (spec[0, 0].real / N, float(np.mean(field ** 2)), m1, m2, spec[fit.ky, fit.kx])


@pytest.mark.parametrize("name", ["triangle", "honeycomb", "stripe"])
@pytest.mark.parametrize("model", ["std", "log"])
def test_spectral_fit_equals_the_real_space_fitness(model, name):
    # the Parseval form of the fitness (what the device diagnostics feed)
    # against the real-space definition, with noise, and for logPFC with
    # amplitudes large enough that the 3-mode rebuild is rescaled
    ad = STD if model == "std" else LOG
    s = ad.STRUCTURES[name]
    rel = model == "log"
    fit = A.spectral_fit(s, CELL, relative=rel, nonnegative=rel)
    rng = np.random.default_rng(5)
    fields = []
    for amp in (0.05, 0.4, 1.5):
        mean = 0.6 if rel else -0.2
        base = A.three_mode(s, CELL, mean, amp * 0.2 * s.amplitudeSign, amp * 0.03, amp * 0.01)
        fields.append(base + 0.02 * rng.standard_normal(base.shape))
    fitness, amps = fit.assemble([_device_rows(f, fit) for f in fields])
    rescaled = 0
    for f, fi, am in zip(fields, fitness, amps):
        ref = A.fitness(f, s, CELL, relative=rel, nonnegative=rel)
        assert fi == pytest.approx(ref[0], abs=1e-13)
        assert tuple(am) == pytest.approx(ref[1:], abs=1e-14)
        if rel:
            a = np.array(ref[1:]) * f.mean() * s.amplitudeSign
            rescaled += A.three_mode(s, CELL, f.mean(), *a).min() < 0
    if rel and name == "triangle":
        assert rescaled >= 1


def test_frozen_moments_from_device_style_sums():
    # pfc_diagnose's m1, m2 (full-spectrum sums with |k|^2, |k|^4) give the
    # FrozenField moments A1, A2 when taken on the q00 = 1 box
    phi = LOG.initial_field("triangle", CELL, {"mode": "x", "A1": 0.9, "A2": 0.1, "A3": 0.0}, 0.7)
    phi = phi + 0.01 * np.random.default_rng(3).standard_normal(phi.shape)
    fit = A.spectral_fit(LOG.STRUCTURES["triangle"], CELL, True, True)
    mean, m0, m1, m2, _ = _device_rows(phi, fit, A.k2_ref(CELL))
    ff = LOG.frozen_field(phi, dict(beta=0.5, g=-2.0), -0.9, CELL)
    assert (m0, m1, m2) == pytest.approx(tuple(ff.A), rel=1e-13)
