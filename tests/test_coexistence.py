"""Common-tangent consumer (host code): pinned on the reference's stored
logPFC triangle table and its tangent, and on closed-form parabola tangents."""
import json
import os

import numpy as np
import pytest

from pfc_hip.coexistence import analytic_branch, common_tangent, std_liquid, std_sweep_coexistence, tabulated_branch

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "common_tangent_triangle_eps_m0p9.json")))


def _log_liquid(lo, hi):
    b, g, e = GOLD["beta"], GOLD["g"], GOLD["eps"]
    return analytic_branch(
        "liquid",
        lambda p: np.asarray(p) * np.log(p) + 0.5 * (b + e) * np.asarray(p) ** 2 + g / 3 * np.asarray(p) ** 3
        + 0.25 * np.asarray(p) ** 4,
        lambda p: 1 + np.log(p) + (b + e) * np.asarray(p) + g * np.asarray(p) ** 2 + np.asarray(p) ** 3,
        lo, hi, d2f=lambda p: 1 / np.asarray(p) + (b + e) + 2 * g * np.asarray(p) + 3 * np.asarray(p) ** 2)


def test_reference_triangle_table_tangent():
    tri = tabulated_branch("triangle", GOLD["phi"], GOLD["fmin"])
    (lo, hi), = tri.nonconvex
    assert lo == GOLD["phi"][0] and abs(hi - GOLD["nonconvex_low_edge_hi"]) < 1e-5
    t = common_tangent(_log_liquid(1e-3, 0.52), tri)
    assert t.status == "ok"
    assert abs(t.phi_a - GOLD["phi_l"]) <= GOLD["tolerance_phi"]
    assert abs(t.phi_b - GOLD["phi_s"]) <= GOLD["tolerance_phi"]
    assert abs(t.mu - GOLD["mu"]) <= GOLD["tolerance_mu_omega"]
    assert abs(t.omega - GOLD["omega"]) <= GOLD["tolerance_mu_omega"]
    # the two tangency conditions hold at the solution
    liq = _log_liquid(1e-3, 0.52)
    assert abs(float(liq.mu(t.phi_a)) - float(tri.mu(t.phi_b))) < 1e-12
    om_a = float(liq.f(t.phi_a)) - t.mu * t.phi_a
    om_b = float(tri.f(t.phi_b)) - t.mu * t.phi_b
    assert abs(om_a - om_b) < 1e-13
    s = common_tangent(tri, liq)
    assert abs(s.phi_a - t.phi_b) < 1e-13 and abs(s.phi_b - t.phi_a) < 1e-13


def _parabola(k, c, e, lo, hi, tabulated):
    if tabulated:
        x = np.linspace(lo, hi, 41)
        return tabulated_branch("p", x, 0.5 * k * (x - c) ** 2 + e)
    return analytic_branch("p", lambda p: 0.5 * k * (np.asarray(p) - c) ** 2 + e, lambda p: k * (np.asarray(p) - c),
                           lo, hi, d2f=lambda p: np.full(np.shape(p), float(k)))


@pytest.mark.parametrize("tabulated", [False, True])
def test_parabola_tangent_is_exact(tabulated):
    # omega_i(mu) = -mu^2/(2 k_i) - c_i mu + e_i; equal omegas give a quadratic in mu
    A, B = (4.0, 0.85, -0.66), (6.0, 1.20, -0.75)
    qa = -(0.5 / A[0] - 0.5 / B[0])
    roots = np.roots([qa, -(A[1] - B[1]), A[2] - B[2]]).real
    exact = [(A[1] + m / A[0], B[1] + m / B[0], m) for m in roots if 0.65 < A[1] + m / A[0] < 1.05]
    (pa, pb, m), = exact
    t = common_tangent(_parabola(*A, 0.65, 1.05, tabulated), _parabola(*B, 0.95, 1.45, tabulated))
    assert t.status == "ok"
    assert abs(t.phi_a - pa) < 1e-10 and abs(t.phi_b - pb) < 1e-10 and abs(t.mu - m) < 1e-9


def test_no_tangent_statuses():
    a = _parabola(4.0, 0.0, 0.0, -0.2, 0.2, False)          # mu in [-0.8, 0.8]
    b = _parabola(4.0, 5.0, 0.0, 5.5, 5.9, False)           # mu in [2.0, 3.6]
    assert common_tangent(a, b).status == "no_overlap"
    c = _parabola(4.0, 0.0, 1.0, -0.2, 0.2, False)          # a shifted up: omega differs by 1 everywhere
    assert common_tangent(a, c).status == "no_root"
    # equal curvature, minima at the same height: f = 0 touches both
    d = _parabola(4.0, 5.0, 0.0, 4.8, 5.2, False)
    t = common_tangent(a, d)
    assert t.status == "ok" and abs(t.phi_a) < 1e-12 and abs(t.phi_b - 5.0) < 1e-12


def test_std_liquid_and_sweep_coexistence_on_a_model_crystal():
    # a crystal branch f_c = f_l(psi) - d(psi): the construction must return
    # densities where both tangency conditions hold
    r = -0.3
    liq = std_liquid(r)
    psi = np.linspace(-0.5, -0.1, 11)
    fc = np.asarray(liq.f(psi)) - 0.02 * np.exp(-((psi + 0.3) / 0.12) ** 2)
    t = std_sweep_coexistence(r, psi, fc)
    assert t.status == "ok"
    tri = tabulated_branch("c", psi, fc)
    assert abs(float(liq.mu(t.phi_a)) - float(tri.mu(t.phi_b))) < 1e-12 or \
        abs(float(tri.mu(t.phi_a)) - float(liq.mu(t.phi_b))) < 1e-12


def test_sweep_maxwell_lines_from_relaxed_states(tmp_path):
    """Maxwell lines of a sweep-shaped table (16 psibar per r) whose crystal
    branch crosses the liquid at psibar = -0.45 with a kink, f_c = f_l -
    A (psibar + 0.45) above it, A growing with -r: every r gives one stable
    liquid-crystal row that equals the pairwise tangent, and the CSV holds
    exactly the rows."""
    import csv
    from pfc_hip.sweep import Instance, maxwell_lines_by_r
    rs = [-0.4, -0.3, -0.2]
    psis = np.linspace(-0.6, 0.0, 16)
    inst, fe = [], []
    for r in rs:
        liq = std_liquid(r)
        for p in psis:
            inst.append(Instance(index=len(inst), r=r, psibar=p, N=64, Lx=1.0, Ly=1.0, seed=0))
            fe.append(float(liq.f(p)) - 0.2 * (-r) * max(p + 0.45, 0.0))
    ml = maxwell_lines_by_r(inst, np.array(fe))
    assert ml.temps == tuple(rs)
    assert [(row["temp"], row["a"], row["b"]) for row in ml.rows] == [(r, "liquid", "crystal") for r in rs]
    for r in rs:
        row = next(x for x in ml.rows if x["temp"] == r)
        t = std_sweep_coexistence(r, psis, [f for i, f in zip(inst, fe) if i.r == r])
        assert t.status == "ok" and row["residual"] < 1e-12
        assert abs(row["phi_a"] - t.phi_a) < 1e-12 and abs(row["phi_b"] - t.phi_b) < 1e-12
    path = tmp_path / "maxwell.csv"
    ml.write_csv(str(path))
    got = list(csv.DictReader(open(path)))
    assert len(got) == len(ml.rows) and float(got[0]["mu"]) == ml.rows[0]["mu"]


def test_reversed_and_degenerate_liquid_crystal_pairs_are_spurious():
    """The sweep's hygiene rule: the liquid must be the lower density by at
    least the gap; a reversed or sliver pair is not a coexistence."""
    from pfc_hip.coexistence import liquid_crystal_pair_ok
    from pfc_hip.phase_diagram.common_tangent import TangentResult
    good = TangentResult("liquid", "crystal", "ok", phi_a=-0.51, phi_b=-0.39)
    reversed_ = TangentResult("liquid", "crystal", "ok", phi_a=-0.4725, phi_b=-0.4769)
    sliver = TangentResult("liquid", "crystal", "ok", phi_a=-0.5818, phi_b=-0.5812)
    none = TangentResult("liquid", "crystal", "no_root")
    assert liquid_crystal_pair_ok(good)
    assert not liquid_crystal_pair_ok(reversed_) and not liquid_crystal_pair_ok(sliver)
    assert not liquid_crystal_pair_ok(none)
