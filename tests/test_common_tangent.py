"""Common tangents, envelopes and Maxwell lines (pfc_hip.phase_diagram.common_tangent, CPU).

Pins:
* the fixture tests/golden/common_tangent_triangle_eps_m0p9.json: the
  reference's stored eps = -0.9 logPFC triangle table (beta = 0.5, g = -2)
  and the coexisting densities recorded for it against the analytic liquid
  (bound 3e-9, the fixture's stated precision), the fsolve solution and the
  notebook's spline-liquid values (1e-13: same construction, round-off);
* closed forms: parabolas f = k/2 (phi - c)^2 + e have omega(mu) = e - c mu
  - mu^2 / (2k), so a tangent is a root of a quadratic in mu, and a
  not-a-knot spline of a parabola is exact;
* invariances a wrong sign, index or operand order would break: adding a
  line alpha phi + beta to every branch shifts mu by alpha and omega by beta
  and leaves the densities; mirroring phi -> c - phi mirrors the densities,
  negates mu and adds c mu to omega; swapping the arguments swaps the sides;
* the envelope does not depend on the tabulation grid.
"""
import csv
import io
import json
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.optimize import fsolve

from pfc_hip.phase_diagram.common_tangent import (STATUSES, analytic_branch, common_tangent, lower_envelope,
                                                  maxwell_lines, tabulated_branch)

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "common_tangent_triangle_eps_m0p9.json")))
BETA, G, EPS = GOLD["beta"], GOLD["g"], GOLD["eps"]


def log_liquid(phi, eps=EPS, shift=(0.0, 0.0)):
    """f, mu, f'' of the uniform logPFC state (+ an optional line alpha phi + beta)."""
    p = np.asarray(phi, dtype=float)
    al, be = shift
    f = p * np.log(p) + 0.5 * (BETA + eps) * p**2 + G / 3 * p**3 + 0.25 * p**4 + al * p + be
    mu = 1 + np.log(p) + (BETA + eps) * p + G * p**2 + p**3 + al
    d2 = 1 / p + (BETA + eps) + 2 * G * p + 3 * p**2
    return f, mu, d2


def liquid(lo, hi, d2=False, name="liquid", shift=(0.0, 0.0)):
    return analytic_branch(name, lambda p: log_liquid(p, shift=shift)[0], lambda p: log_liquid(p, shift=shift)[1],
                           lo, hi, d2f=(lambda p: log_liquid(p)[2]) if d2 else None)


def triangle(shift=(0.0, 0.0)):
    phi, f, q = (np.asarray(GOLD[k], dtype=float) for k in ("phi", "fmin", "q0star"))
    return tabulated_branch("triangle", phi, f + shift[0] * phi + shift[1], q)


def parabola(name, k, c, e, lo, hi, table=False, q=None):
    fn = lambda p: 0.5 * k * (np.asarray(p, dtype=float) - c) ** 2 + e  # noqa: E731
    if table:
        x = np.linspace(lo, hi, 41)
        return tabulated_branch(name, x, fn(x), None if q is None else q(x))
    return analytic_branch(name, fn, lambda p: k * (np.asarray(p, dtype=float) - c), lo, hi,
                           d2f=lambda p: np.full(np.shape(p), float(k)))


def tangent_slopes(pa, pb):
    """Slopes mu of the tangents of parabolas pa = (k, c, e) and pb: roots of
    omega_a(mu) - omega_b(mu) = (ea - eb) - (ca - cb) mu - (1/ka - 1/kb) mu^2 / 2."""
    (ka, ca, ea), (kb, cb, eb) = pa, pb
    quad = -0.5 * (1 / ka - 1 / kb)
    if quad == 0:
        return [(ea - eb) / (ca - cb)]
    return sorted(np.roots([quad, -(ca - cb), ea - eb]).real.tolist())


def close(a, b, tol):
    return all(abs(x - y) <= tol for x, y in zip(a, b))


# -- the stored table ---------------------------------------------------------
def test_triangle_table_pieces():
    tri = triangle()
    assert len(tri.segments) == 1 and not tri.gaps and not tri.dropped
    (lo, hi), = tri.nonconvex
    assert lo == GOLD["phi"][0] and abs(hi - GOLD["nonconvex_low_edge_hi"]) <= 1e-5
    (p,) = tri.pieces
    assert p.lo == hi and p.hi == GOLD["phi"][-1]
    # f'' of the spline changes sign at hi: negative before, positive after
    d2 = tri.segments[0].d2f
    assert d2(hi - 1e-4) < 0 < d2(hi + 1e-4)


def test_recorded_liquid_triangle_coexistence():
    t = common_tangent(liquid(1e-3, 0.52), triangle())
    assert t.ok and t.supporting and not t.same_branch
    assert abs(t.phi_a - GOLD["phi_l"]) <= 3e-9 and abs(t.phi_b - GOLD["phi_s"]) <= 3e-9
    assert abs(t.mu - GOLD["mu"]) <= 1e-6 and abs(t.omega - GOLD["omega"]) <= 1e-6
    fs = GOLD["fsolve_p0201"]
    assert abs(t.phi_a - fs["phi_l"]) < 1e-9 and abs(t.phi_b - fs["phi_s"]) < 1e-9
    assert t.q0star_a is None and abs(t.q0star_b - fs["q"]) < 1e-9
    assert t.residual < 1e-14
    # the tangent conditions, evaluated independently
    f_l, mu_l, _ = log_liquid(t.phi_a)
    tri = triangle()
    assert abs(mu_l - t.mu) < 1e-13 and abs(float(tri.mu(t.phi_b)) - t.mu) < 1e-12
    assert abs((float(tri.f(t.phi_b)) - f_l) / (t.phi_b - t.phi_a) - t.mu) < 1e-12


def test_notebook_spline_liquid_gives_the_recorded_densities():
    """The recorded densities came from splines of the liquid's f and mu on a
    0.01 grid: that representation reproduces them to round-off."""
    x = np.linspace(0.01, 0.52, 51, endpoint=False)
    liq = analytic_branch("liquid", CubicSpline(x, log_liquid(x)[0]), CubicSpline(x, log_liquid(x)[1]), x[0], x[-1])
    t = common_tangent(liq, triangle())
    assert t.ok
    assert abs(t.phi_a - GOLD["phi_l"]) < 1e-13 and abs(t.phi_b - GOLD["phi_s"]) < 1e-13
    assert abs(t.q0star_b - GOLD["notebook_q_at_ref_phi_s"]) < 1e-13


def test_argument_swap_and_liquid_domain():
    liq, tri = liquid(1e-3, 0.52), triangle()
    t, s = common_tangent(liq, tri), common_tangent(tri, liq)
    assert s.ok and close((s.phi_b, s.phi_a, s.mu, s.omega), (t.phi_a, t.phi_b, t.mu, t.omega), 1e-13)
    wide = liquid(1e-3, 1.6)  # across its spinodal region: two convex pieces
    assert len(wide.pieces) == 2 and len(wide.nonconvex) == 1
    u = common_tangent(wide, tri)
    assert u.ok and close((u.phi_a, u.phi_b), (t.phi_a, t.phi_b), 1e-13)


@pytest.mark.parametrize("alpha,beta", [(0.3, -0.2), (-1.1, 0.05)])
def test_adding_a_line_shifts_mu_and_omega_only(alpha, beta):
    t = common_tangent(liquid(1e-3, 0.52), triangle())
    u = common_tangent(liquid(1e-3, 0.52, shift=(alpha, beta)), triangle(shift=(alpha, beta)))
    assert u.ok and close((u.phi_a, u.phi_b), (t.phi_a, t.phi_b), 1e-11)
    assert abs(u.mu - (t.mu + alpha)) < 1e-11 and abs(u.omega - (t.omega + beta)) < 1e-11


def test_envelope_of_liquid_and_triangle_is_their_tangent():
    liq, tri = liquid(1e-3, 0.52), triangle()
    env = lower_envelope([liq, tri], temp=EPS)
    assert [w.name for w in env.windows] == ["liquid", "triangle"]
    (c,) = env.coexistence
    t = common_tangent(liq, tri)
    assert c.stable and (c.a, c.b) == ("liquid", "triangle")
    assert close((c.phi_a, c.phi_b, c.mu, c.omega), (t.phi_a, t.phi_b, t.mu, t.omega), 1e-14)
    assert env.windows[0].phi_hi == c.phi_a and env.windows[1].phi_lo == c.phi_b
    assert env.metastable == () and env.warnings == ()


# -- parabolas ----------------------------------------------------------------
P1, P2 = (4.0, 0.85, -0.66), (6.0, 1.20, -0.75)


@pytest.mark.parametrize("table", [False, True])
def test_parabola_tangent_closed_form(table):
    q1 = lambda x: 0.98 - 0.1 * (x - 0.85)  # noqa: E731
    a = parabola("stripe", *P1, 0.65, 1.05, table, q=q1 if table else None)
    b = parabola("honeycomb", *P2, 0.95, 1.45, table)
    m = [s for s in tangent_slopes(P1, P2) if 0.65 < P1[1] + s / P1[0] < 1.05]
    assert len(m) == 1
    m = m[0]
    t = common_tangent(a, b)
    assert t.ok and len(t.candidates) == 1
    want = (P1[1] + m / P1[0], P2[1] + m / P2[0], m, P1[2] - P1[1] * m - m * m / (2 * P1[0]))
    assert close((t.phi_a, t.phi_b, t.mu, t.omega), want, 1e-14)
    if table:
        assert abs(t.q0star_a - q1(t.phi_a)) < 1e-14 and t.q0star_b is None


def test_denser_liquid_and_mirror_image():
    hc, lq = (6.0, 1.10, -0.70), (2.0, 1.60, -0.72)
    a, b = parabola("honeycomb", *hc, 0.90, 1.35), parabola("liquid", *lq, 1.0, 2.2)
    t = common_tangent(a, b)
    assert t.ok and t.phi_b > t.phi_a  # the liquid is the denser phase
    (m,) = [s for s in tangent_slopes(hc, lq) if 0.9 < hc[1] + s / hc[0] < 1.35]
    assert close((t.phi_a, t.phi_b, t.mu), (hc[1] + m / hc[0], lq[1] + m / lq[0], m), 1e-14)
    c = 3.0
    am = parabola("honeycomb", hc[0], c - hc[1], hc[2], c - 1.35, c - 0.90)
    bm = parabola("liquid", lq[0], c - lq[1], lq[2], c - 2.2, c - 1.0)
    r = common_tangent(bm, am)
    assert r.ok and r.phi_a < r.phi_b
    assert close((r.phi_b, r.phi_a), (c - t.phi_a, c - t.phi_b), 1e-14)
    assert abs(r.mu + t.mu) < 1e-14 and abs(r.omega - (t.omega + c * t.mu)) < 1e-14


def test_two_tangents_and_candidate_order():
    """Parabolas of different curvature whose mu ranges overlap widely have
    two tangents; the primary is the supporting one, `ordered` and
    `mu_guess` select among them."""
    pa, pb = (1.0, 0.0, 0.0), (9.0, 0.5, 0.01)
    a, b = parabola("a", *pa, -2.0, 2.0), parabola("b", *pb, 0.2, 0.8)
    slopes = [s for s in tangent_slopes(pa, pb) if -2 < s / pa[0] < 2 and 0.2 < pb[1] + s / pb[0] < 0.8]
    t = common_tangent(a, b)
    assert t.ok and len(t.candidates) == len(slopes) == 2
    assert sorted(c.mu for c in t.candidates) == pytest.approx(slopes, abs=1e-13)
    # b lies above a's tangent lines: both tangents support the pair; the
    # primary is the lowest mu, or the one nearest mu_guess
    assert all(c.supporting for c in t.candidates)
    assert abs(t.mu - min(slopes)) < 1e-13
    assert abs(common_tangent(a, b, mu_guess=max(slopes)).mu - max(slopes)) < 1e-13
    o = common_tangent(a, b, ordered=True)
    assert o.ok and all(c.phi_a < c.phi_b for c in o.candidates) and len(o.candidates) == 1


# -- pairs without a tangent --------------------------------------------------
def test_status_no_overlap():
    r = common_tangent(parabola("a", 1.0, 0.0, 0.0, -0.1, 0.1), parabola("b", 1.0, 2.0, 0.0, 1.5, 1.7))
    assert r.status == "no_overlap"


def test_status_no_root_for_parallel_curves_and_a_convex_self_pair():
    a, b = parabola("a", 1.0, 1.0, 0.0, 0.8, 1.2), parabola("b", 1.0, 1.0, 0.1, 0.7, 1.3)
    assert common_tangent(a, b).status == "no_root"
    assert common_tangent(a, a).status == "no_root"


def test_status_non_convex():
    concave = analytic_branch("c", lambda p: -np.asarray(p) ** 2, lambda p: -2 * np.asarray(p), 0.0, 1.0)
    assert concave.pieces == () and concave.nonconvex == ((0.0, 1.0),)
    assert common_tangent(concave, parabola("b", 1.0, 1.0, 0.1, 0.7, 1.3)).status == "non_convex"
    # f'' = 1 - phi: convex up to phi = 1 (mu in [0, 0.5]); a partner whose
    # slopes lie in the concave part's range only
    cubic = analytic_branch("q", lambda p: np.asarray(p) ** 2 / 2 - np.asarray(p) ** 3 / 6,
                            lambda p: np.asarray(p) - np.asarray(p) ** 2 / 2, 0.0, 3.0)
    assert abs(cubic.pieces[0].hi - 1.0) < 1e-10
    r = common_tangent(cubic, parabola("b", 1.0, 5.0, 0.0, 3.8, 4.5))
    assert r.status == "non_convex"
    for res in (r,):
        assert res.status in STATUSES and not res.ok and np.isnan(res.phi_a) and res.message


# -- envelopes ----------------------------------------------------------------
PA, PB, PC = (4.0, 0.3, 0.0), (4.0, 0.6, -0.05), (4.0, 0.9, 0.0)


def three(eb=PB[2], kb=PB[0], dom_b=(0.35, 0.85)):
    return [parabola("A", *PA, 0.0, 0.5), parabola("B", kb, PB[1], eb, *dom_b), parabola("C", *PC, 0.7, 1.2)]


def test_three_phase_envelope():
    env = lower_envelope(three())
    assert [w.name for w in env.windows] == ["A", "B", "C"]
    for got, (pl, pr) in zip(env.coexistence, [(PA, PB), (PB, PC)]):
        (m,) = tangent_slopes(pl, pr)
        assert got.stable and got.ok and close((got.mu, got.phi_a, got.phi_b),
                                               (m, pl[1] + m / pl[0], pr[1] + m / pr[0]), 1e-14)
    assert env.windows[1].phi_lo == env.coexistence[0].phi_b and env.windows[1].phi_hi == env.coexistence[1].phi_a
    (ac,) = env.metastable
    assert (ac.a, ac.b, ac.stable, ac.below) == ("A", "C", False, "B") and abs(ac.mu) < 1e-14
    assert env.warnings == ()
    # the tabulated envelope lies on or below every branch, and on the windows' branches
    for br in three():
        f = br.f(env.hull_phi)
        ok = np.isfinite(f)
        assert (env.hull_f[ok] <= f[ok] + 1e-12).all()


@pytest.mark.parametrize("grid", [None, np.linspace(0.0, 1.2, 7), [0.0, 0.6, 1.2]])
def test_envelope_does_not_depend_on_the_grid(grid):
    """A narrow deep B (k = 400) is found whatever the grid; a narrow B above
    the A-C tangent never enters."""
    ref = lower_envelope(three(eb=-0.001, kb=400.0))
    env = lower_envelope(three(eb=-0.001, kb=400.0), grid)
    assert [w.name for w in env.windows] == ["A", "B", "C"]
    assert [(c.phi_a, c.phi_b, c.mu) for c in env.coexistence] == [(c.phi_a, c.phi_b, c.mu) for c in ref.coexistence]
    hi = lower_envelope(three(eb=0.0005, kb=400.0, dom_b=(0.55, 0.65)), grid)
    assert [w.name for w in hi.windows] == ["A", "C"]
    (ac,) = hi.coexistence
    assert abs(ac.mu) < 1e-14 and abs(ac.phi_a - 0.3) < 1e-14 and abs(ac.phi_b - 0.9) < 1e-14
    assert {(m.a, m.b) for m in hi.metastable} == {("A", "B"), ("B", "C")}


def test_liquid_liquid_coexistence_across_the_spinodal():
    exact = liquid(1e-3, 1.6, d2=True)
    numeric = liquid(1e-3, 1.6)
    spin = np.sort(np.roots([3.0, 2.0 * G, BETA + EPS, 1.0]).real)[1:]  # phi f'' = 0, positive roots
    assert close([exact.pieces[0].hi, exact.pieces[1].lo], spin, 1e-14)
    assert close([numeric.pieces[0].hi, numeric.pieces[1].lo], spin, 1e-9)
    env = lower_envelope([exact])
    (ll,) = env.coexistence
    assert ll.same_branch and not ll.across_gap and ll.stable and ll.supporting
    assert ll.phi_a < spin[0] < spin[1] < ll.phi_b and ll.residual < 1e-14
    ref = fsolve(lambda v: [log_liquid(v[0])[1] - log_liquid(v[1])[1],
                            log_liquid(v[1])[0] - log_liquid(v[0])[0] - log_liquid(v[0])[1] * (v[1] - v[0])],
                 [0.35, 1.45], xtol=1e-13)
    assert close((ll.phi_a, ll.phi_b), ref, 1e-10)
    # with the triangle present the liquid-liquid tangent is metastable, under the triangle
    env2 = lower_envelope([numeric, triangle()])
    (meta,) = [m for m in env2.metastable if m.same_branch]
    assert meta.below == "triangle" and abs(meta.phi_a - ll.phi_a) < 1e-9


def test_nan_runs_become_segments_and_gap_bridges_are_flagged():
    x = np.round(np.linspace(-1.5, 1.5, 61), 12)
    f = (x**2 - 1) ** 2
    f[np.abs(x) < 0.4] = np.nan
    f[2] = np.nan  # leaves entries 0 and 1: a run too short for a spline
    dw = tabulated_branch("dw", x, f)
    assert [(s.lo, s.hi) for s in dw.segments] == [(-1.35, -0.4), (0.4, 1.5)]
    assert dw.gaps == ((-0.4, 0.4),) and dw.dropped == (-1.5, -1.45)
    assert [p.segment for p in dw.pieces] == [0, 1] and len(dw.nonconvex) == 2
    t = common_tangent(dw, dw)
    assert t.ok and t.same_branch and t.across_gap
    assert abs(t.phi_a + 1) < 1e-7 and abs(t.phi_b - 1) < 1e-7 and abs(t.mu) < 1e-12
    env = lower_envelope([dw])
    assert env.coexistence[0].across_gap and any("NaN gap" in w for w in env.warnings)


# -- Maxwell lines --------------------------------------------------------------
def test_triple_point_of_a_closing_window():
    """e_B(T) = -0.05 + 0.11 T: B's window closes at T* = 0.05/0.11, where
    the A-B and B-C tangents meet the A-C one; the width of B's window is
    linear in T for equal curvatures, so the estimate is exact."""
    temps = [0.0, 0.15, 0.3, 0.45, 0.6, 0.75]
    envs = {t: lower_envelope(three(eb=-0.05 + 0.11 * t), temp=t) for t in temps}
    ml = maxwell_lines(envs)
    assert ml.temps == tuple(temps)
    want = []
    for t in temps:
        want += [(t, "A", "B"), (t, "B", "C")] if t < 0.05 / 0.11 else [(t, "A", "C")]
    assert [(r["temp"], r["a"], r["b"]) for r in ml.rows] == want
    (tp,) = ml.triple_points
    assert tp.phases == ("A", "B", "C") and (tp.temp_open, tp.temp_closed) == (0.45, 0.6)
    assert tp.width_open > 0 > tp.width_closed
    assert abs(tp.temp_estimate - 0.05 / 0.11) < 1e-12
    assert maxwell_lines(list(envs.values())).triple_points == ml.triple_points
    buf = io.StringIO()
    ml.write_csv(buf)
    back = list(csv.DictReader(io.StringIO(buf.getvalue())))
    assert list(back[0]) == ["temp", "a", "b", "phi_a", "phi_b", "mu", "omega", "q0star_a", "q0star_b", "status",
                             "same_branch", "across_gap", "stable", "below", "residual"]
    assert len(back) == len(ml.rows) and all(float(b["phi_b"]) == r["phi_b"] for b, r in zip(back, ml.rows))
    w = io.StringIO()
    ml.write_csv(w, "windows")
    assert w.getvalue().splitlines()[0] == "temp,name,phi_lo,phi_hi"
