"""The phase-diagram driver over the hip adapters (needs a GPU): logPFC
evaluations and the q00 scan against the values the reference's driver tests
hold (tests/test_phase_diagram_q0star.py:44-64, 486-556, 612-643), the
uniform state, the stdPFC g = 0 mirror, and `run` against a step loop.

Tolerances are the reference's own for these comparisons: f to 3e-14 and
(fitness, A, B, C) to 1e-12 per evaluation (2e-14 and 1e-13 without ansatz
noise), q0* to 1e-8 and fmin to 1e-13, the honeycomb's converged f to 1e-12.
"""
import dataclasses
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REFS = json.load(open(os.path.join(HERE, "golden", "phase_diagram_refs.json")))


def _settings():
    from pfc_hip.phase_diagram import q0star as Q
    return Q.RelaxSettings(**REFS["relax"]), Q.QScanSettings(**REFS["qscan"])


def _log_minimizer(structure, **relax_changes):
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.phase_diagram import q0star as Q
    relax, qscan = _settings()
    relax = dataclasses.replace(relax, **relax_changes)
    return Q.MinimizeFreeEnergy(LOG, REFS["log_params"], structure, REFS["cell"], REFS["ansatz"], relax, qscan)


def _ref(structure):
    return next(r for r in REFS["get_f_abc"] if r["structure"] == structure)


@pytest.mark.parametrize("structure", ["triangle", "stripe"])
def test_log_get_f_abc_against_reference(hip_device, structure):
    from pfc_hip.phase_diagram import q0star as Q
    r = _ref(structure)
    m = _log_minimizer(structure)
    f, fit, A, B, C = (float(x) for x in m.Get_f_ABC(temp=r["eps"], meanDensity=r["phi"], q00=r["q00"]))
    print(f"\n{structure}: df {f - r['f']:.3e} dfit {fit - r['fitness']:.3e} dA {A - r['A']:.3e} "
          f"dB {B - r['B']:.3e} dC {C - r['C']:.3e} steps {m.last_steps}", flush=True)
    assert m.last_stop == Q.STOP_CONVERGED
    assert m.last_steps == r["steps"]
    assert abs(f - r["f"]) <= r["tol_f"]
    assert (fit, A, B, C) == pytest.approx((r["fitness"], r["A"], r["B"], r["C"]), abs=r["tol_diag"])


@pytest.mark.parametrize("structure", ["triangle", "stripe"])
def test_log_get_f_abc_without_noise(hip_device, structure):
    r = _ref(structure)
    m = _log_minimizer(structure, noise=0.0)
    f, fit, A, B, C = (float(x) for x in m.Get_f_ABC(temp=r["eps"], meanDensity=r["phi"], q00=r["q00"]))
    print(f"\n{structure} noise 0: df {f - r['f']:.3e} dfit {fit - r['fitness']:.3e} dA {A - r['A']:.3e} "
          f"steps {m.last_steps}", flush=True)
    assert m.last_steps == r["steps"]
    assert abs(f - r["f"]) <= 2e-14
    assert (fit, A, B, C) == pytest.approx((r["fitness"], r["A"], r["B"], r["C"]), abs=1e-13)


def test_log_honeycomb_integrates_through_the_transient(hip_device):
    # dt = 0.1 drives phi negative in the first block (NaN from the log): the
    # rollback redoes it at reduced dt and the evaluation converges on the
    # relaxed honeycomb, far below the transient the unrefined run stopped at
    from pfc_hip.phase_diagram import q0star as Q
    r = _ref("honeycomb")
    m = _log_minimizer("honeycomb")
    f = float(m.Get_f_ABC(temp=r["eps"], meanDensity=r["phi"], q00=r["q00"])[0])
    print(f"\nhoneycomb: df {f - r['f']:.3e} steps {m.last_steps}", flush=True)
    assert m.last_stop == Q.STOP_CONVERGED
    assert abs(f - r["f"]) <= r["tol_f"]
    assert f < r["f_original_transient"] - r["below_original_by"]
    assert m.last_steps == r["steps"]


def test_log_find_q0star_against_reference(hip_device):
    ref = REFS["find_q0star"]
    recs = []
    q, f = _log_minimizer(ref["structure"]).Find_q0Star(temp=ref["eps"], meanDensity=ref["phi"],
                                                        record=recs.append)
    print(f"\nq0*: dq {float(q) - ref['q0star']:.3e} dfmin {float(f) - ref['fmin']:.3e} "
          f"evaluations {len(recs)} steps {sum(r['steps'] for r in recs)}", flush=True)
    assert abs(float(q) - ref["q0star"]) <= ref["tol_q"]
    assert abs(float(f) - ref["fmin"]) <= ref["tol_f"]


def test_log_points_without_a_local_minimum(hip_device):
    from pfc_hip.phase_diagram import q0star as Q
    ref = REFS["no_local_minimum"]
    r = Q.scan_density(_log_minimizer(ref["structure"]), ref["eps"], ref["phi"])
    for k in range(len(ref["phi"])):
        assert (str(r["status"][k]), str(r["message"][k])) == (Q.STATUS_NO_LOCAL_MINIMUM, ref["message"])
        assert np.isnan(r["q0star"][k]) and np.isnan(r["fmin"][k])


@pytest.mark.parametrize("slug, density", [("logpfc_hip", 0.7), ("stdpfc_hip", -0.3)])
def test_liquid_free_energy_is_the_uniform_state(hip_device, slug, density):
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.models.stdpfc_hip import phase_diagram as STD
    ad = {"logpfc_hip": LOG, "stdpfc_hip": STD}[slug]
    params, eps = dict(beta=0.5, g=-2.0), -0.9
    sim = ad.build(params, "triangle", REFS["cell"], REFS["ansatz"], eps, density, 1.0, 0.1,
                   state=np.full((24, 28), density))
    assert float(ad.free_energy(sim)) == pytest.approx(ad.liquid_free_energy(density, params, eps), rel=1e-14)


def test_std_g0_triangle_mirrors_honeycomb(hip_device):
    # at g = 0 f(-n) = f(n) and the dynamics are odd in n: the triangle at
    # -nbar is the exact negative of the honeycomb at +nbar, step for step
    from pfc_hip.models.stdpfc_hip import phase_diagram as STD
    from pfc_hip.phase_diagram import q0star as Q
    ref = REFS["std_mirror"]
    relax, qscan = _settings()
    tri = Q.MinimizeFreeEnergy(STD, ref["params"], "triangle", REFS["cell"], REFS["ansatz"], relax, qscan)
    hc = Q.MinimizeFreeEnergy(STD, ref["params"], "honeycomb", REFS["cell"], REFS["ansatz"], relax, qscan)
    a = [float(x) for x in tri.Get_f_ABC(temp=ref["eps"], meanDensity=-ref["nbar"], q00=1.0)]
    b = [float(x) for x in hc.Get_f_ABC(temp=ref["eps"], meanDensity=+ref["nbar"], q00=1.0)]
    print(f"\nmirror: df {a[0] - ref['f']:.3e} dfit {a[1] - ref['fitness']:.3e} dA {a[2] - ref['A']:.3e} "
          f"steps {tri.last_steps}", flush=True)
    assert a == b and tri.last_steps == hc.last_steps
    assert a[1] > 0.999 and a[2] > 0.1
    assert a[0] < STD.liquid_free_energy(ref["nbar"], ref["params"], ref["eps"])
    # the reference's values at this point
    assert tri.last_steps == ref["steps"]
    assert abs(a[0] - ref["f"]) <= 1e-13
    assert (a[1], a[2]) == pytest.approx((ref["fitness"], ref["A"]), abs=1e-11)


def test_run_matches_a_step_loop(hip_device):
    # the driver advances a block with the adapter's `run` (one device call);
    # without it, with `step` per time step: the same evaluation
    import types
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.phase_diagram import q0star as Q
    steps_only = types.SimpleNamespace(**{k: getattr(LOG, k) for k in dir(LOG) if not k.startswith("_")})
    del steps_only.run
    relax, qscan = _settings()
    relax = dataclasses.replace(relax, totalT=300.0)
    out = []
    for ad in (LOG, steps_only):
        m = Q.MinimizeFreeEnergy(ad, REFS["log_params"], "triangle", REFS["cell"], REFS["ansatz"], relax, qscan)
        out.append(([float(x) for x in m.Get_f_ABC(temp=-0.9, meanDensity=0.63, q00=0.96)], m.last_steps))
    assert out[0] == out[1]


# -- the batched driver ------------------------------------------------------------
def test_batch_slot_box_matches_a_plan(hip_device):
    # a batch slot given its own box and dt steps like a plan built with them
    from pfc_hip import fields as F
    from pfc_hip.plan import Batch, Plan
    kw = dict(model="log", Nx=28, Ny=24, Gamma=1.0, beta=0.5, q0=1.0, epsilon=-0.9, g=-2.0, precision=64)
    a, h = 4 * np.pi / (3 ** 0.5 * 0.97), 4 * np.pi / 0.97
    phi0 = 0.63 + F.uniform_noise((24, 28), 0.0, 0.02, seed=3)
    pl = Plan(Lx=2 * a, Ly=h, dt=0.05, **kw)
    b = Batch(3, Lx=10.0, Ly=20.0, dt=0.1, **kw)
    b.set_slot_box(1, 2 * a, h, 0.05)
    for s in range(3):
        b.set_field(s, phi0)
    pl.set_field(phi0)
    pl.step(200)
    b.step(200)
    x = pl.get_field().cpu().numpy()
    y = b.get_field(1).cpu().numpy()
    print(f"\nslot vs plan: max |d phi| {np.abs(x - y).max():.3e}", flush=True)
    assert np.abs(x - y).max() <= 1e-14
    assert b.reduce()[1, 0] == pytest.approx(pl.free_energy(), abs=1e-15)
    # the other slots kept the shared box: they differ
    assert np.abs(b.get_field(0).cpu().numpy() - x).max() > 1e-6


def test_batched_pass_equals_serial_evaluations(hip_device):
    from pfc_hip.phase_diagram import batched as B
    from pfc_hip.phase_diagram import q0star as Q
    ref = REFS["find_q0star"]
    m = _log_minimizer(ref["structure"])
    _, qscan = _settings()
    qs = Q.pass_grid(qscan.q00_start, qscan.Deltaq, qscan)
    import time
    t0 = time.perf_counter()
    serial = [m.evaluate(ref["eps"], ref["phi"], q) for q in qs]
    t1 = time.perf_counter()
    batched = B.evaluate_batched(m, [(ref["eps"], ref["phi"], q) for q in qs])
    t2 = time.perf_counter()
    df = max(abs(s.f - b.f) for s, b in zip(serial, batched))
    dfit = max(abs(s.fitness - b.fitness) for s, b in zip(serial, batched))
    print(f"\npass of {len(qs)}: serial {t1 - t0:.2f} s, batched {t2 - t1:.2f} s; max df {df:.3e} dfit {dfit:.3e}; "
          f"steps {sum(s.steps for s in serial)}", flush=True)
    assert [s.stop for s in serial] == [b.stop for b in batched]
    assert [s.steps for s in serial] == [b.steps for b in batched]
    assert df <= 1e-14 and dfit <= 1e-12


def test_batched_honeycomb_rollback(hip_device):
    from pfc_hip.phase_diagram import batched as B
    from pfc_hip.phase_diagram import q0star as Q
    r = _ref("honeycomb")
    ev, = B.evaluate_batched(_log_minimizer("honeycomb"), [(r["eps"], r["phi"], r["q00"])])
    print(f"\nbatched honeycomb: df {ev.f - r['f']:.3e} steps {ev.steps}", flush=True)
    assert ev.stop == Q.STOP_CONVERGED and ev.steps == r["steps"]
    assert abs(ev.f - r["f"]) <= r["tol_f"]


def test_scan_density_batched_against_reference(hip_device):
    import time
    from pfc_hip.phase_diagram import batched as B
    from pfc_hip.phase_diagram import q0star as Q
    ref, nlm = REFS["find_q0star"], REFS["no_local_minimum"]
    rhos = nlm["phi"] + [ref["phi"]]
    t0 = time.perf_counter()
    r = B.scan_density_batched(_log_minimizer(ref["structure"]), ref["eps"], rhos)
    print(f"\nbatched scan of {len(rhos)} densities: {time.perf_counter() - t0:.2f} s, "
          f"{int(r['qscan']['steps'].sum())} steps", flush=True)
    assert list(r["status"]) == [Q.STATUS_NO_LOCAL_MINIMUM] * len(nlm["phi"]) + [Q.STATUS_OK]
    assert all(str(msg) == nlm["message"] for msg in r["message"][:len(nlm["phi"])])
    assert abs(r["q0star"][-1] - ref["q0star"]) <= ref["tol_q"]
    assert abs(r["fmin"][-1] - ref["fmin"]) <= ref["tol_f"]


# -- frozen_q on the hip logPFC adapter (the reference's
#    tests/test_phase_diagram_frozen_q.py:10-83 on this backend) ----------------
SMALL = {"uNx": 7, "uNy": 12, "Mx": 2, "My": 1}
H2L = {"mode": "H2L", "A1": 0.9, "A2": 0.0, "A3": 0.0}


def _frozen_relax():
    from pfc_hip.phase_diagram import q0star as Q
    return Q.RelaxSettings(dt=0.01, dtMax=0.1, dtRate=1e-5, dtGrow=2.0, dtChecks=10, totalT=1e6, stepT=100.0,
                           divs=10, maxDepth=4, requiredFitness=0.99, minimumA=1e-6)


def _f_adapter(phi, params, cell, temp, q):
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    sim = LOG.build(params, "triangle", cell, H2L, temp, float(phi.mean()), q, 0.01, state=np.array(phi))
    return float(LOG.free_energy(sim))


@pytest.mark.parametrize("beta,g,temp", [(0.5, -2.0, -0.9), (0.25, 1.0, -3.0), (1.0, 0.0, -2.0)])
def test_frozen_f_equals_the_hip_free_energy(hip_device, beta, g, temp):
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    params = {"beta": beta, "g": g, "v0": 1.0}
    rng = np.random.default_rng(1)
    sim = LOG.build(params, "triangle", REFS["cell"], H2L, temp, 0.7, 1.0, 0.01)
    phi = LOG.get_state(sim) * (1 + 0.05 * rng.standard_normal(LOG.get_state(sim).shape))
    ff = LOG.frozen_field(phi, params, temp, REFS["cell"])
    for q in (0.7, 0.93, 1.0, 1.08, 1.3):
        fa = _f_adapter(phi, params, REFS["cell"], temp, q)
        assert abs(float(ff.f(q)) - fa) <= 1e-12 * max(1.0, abs(fa))


def test_get_f_abc_restarts_from_a_given_state(hip_device):
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.phase_diagram import q0star as Q
    params = {"beta": 0.5, "g": -2.0, "v0": 1.0}
    _, qscan = _settings()
    m = Q.MinimizeFreeEnergy(LOG, params, "triangle", SMALL, H2L, _frozen_relax(), qscan)
    f, fit, *_ = m.Get_f_ABC(-0.9, 0.62, 1.0)
    phi = m.last_state
    assert phi is not None and abs(phi.mean() - 0.62) < 1e-12
    assert abs(_f_adapter(phi, params, SMALL, -0.9, 1.0) - f) < 1e-14
    f2, *_ = m.Get_f_ABC(-0.9, 0.62, 1.0, state=phi)
    assert abs(f2 - f) <= 1e-11 and m.last_steps < 5000


@pytest.mark.parametrize("phi", [0.62, 0.66])
def test_frozen_q_agrees_with_find_q0star(hip_device, phi):
    import time
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.phase_diagram import frozen_q as FQ
    from pfc_hip.phase_diagram import q0star as Q
    params = {"beta": 0.5, "g": -2.0, "v0": 1.0}
    _, qscan = _settings()
    t0 = time.perf_counter()
    q_f, f_f = Q.MinimizeFreeEnergy(LOG, params, "triangle", SMALL, H2L, _frozen_relax(), qscan).Find_q0Star(-0.9, phi)
    t1 = time.perf_counter()
    m = FQ.MinimizeFrozenQ(LOG, params, "triangle", SMALL, H2L, _frozen_relax())
    q, f = m.Find_q0Star(-0.9, phi)
    t2 = time.perf_counter()
    print(f"\nphi {phi}: scan q* {q_f:.10f} ({t1 - t0:.1f} s), frozen q* {q:.10f} ({t2 - t1:.1f} s, "
          f"{m.last_result['iters']} iterations); dq {q - q_f:.2e} df {f - f_f:.2e}", flush=True)
    assert m.last_result["status"] == "ok"
    assert abs(q - q_f) < 1e-5
    assert abs(f - f_f) < 1e-9
    assert abs(m.last_result["f_frozen"] - f) < 1e-10
    assert "defect_flag" in m.last_result


# -- device diagnostics -------------------------------------------------------------
def test_device_diagnostics_equal_numpy_on_the_field(hip_device):
    from pfc_hip import fields as F
    from pfc_hip.models import _adapter as A
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.plan import Batch, Plan
    cell = REFS["cell"]
    fit = A.spectral_fit(LOG.STRUCTURES["triangle"], cell, True, True)
    qs = (0.97, 1.0, 1.04)
    kw = dict(model="log", Nx=28, Ny=24, Gamma=1.0, beta=0.5, q0=1.0, epsilon=-0.9, g=-2.0, dt=0.1, precision=64)
    base = LOG.initial_field("triangle", cell, H2L, 0.63)
    fields = [base + F.uniform_noise(base.shape, 0.0, 0.01, seed=40 + i) for i in range(len(qs))]
    b = Batch(len(qs), **kw, Lx=1.0, Ly=1.0)
    plans = []
    for i, q in enumerate(qs):
        g = A.box(cell, q)
        b.set_slot_box(i, g.Lx, g.Ly, 0.1)
        b.set_field(i, fields[i])
        plans.append(Plan(**kw, Lx=g.Lx, Ly=g.Ly).set_field(fields[i]).step(50))
    b.step(50)
    got = b.get_fields().cpu().numpy()
    rows = b.diagnose(fit.ky, fit.kx)
    N = 28 * 24
    for i, q in enumerate(qs):
        x = plans[i].get_field().cpu().numpy()
        assert np.array_equal(got[i], x) and np.array_equal(b.get_field(i).cpu().numpy(), x)
        mean, m0, m1, m2, ent = plans[i].diagnose(fit.ky, fit.kx)
        bm = rows[i]
        assert (bm[0], bm[1], bm[2], bm[3]) == (mean, m0, m1, m2) and np.array_equal(bm[4], ent)
        spec = np.fft.rfft2(x)
        full = np.abs(np.fft.fft2(x) / N) ** 2
        k2 = A.k2_ref(cell) * q * q
        assert mean == pytest.approx(spec[0, 0].real / N, abs=1e-15)
        assert m0 == pytest.approx(np.mean(x ** 2), rel=1e-14)
        assert (m1, m2) == pytest.approx(((full * k2).sum(), (full * k2 * k2).sum()), rel=1e-13)
        assert np.abs(ent - spec[fit.ky, fit.kx]).max() <= 1e-12 * N


@pytest.mark.parametrize("slug", ["logpfc_hip", "stdpfc_hip"])
@pytest.mark.parametrize("structure", ["triangle", "honeycomb", "stripe"])
def test_adapter_device_diagnose_equals_host(hip_device, slug, structure):
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.models.stdpfc_hip import phase_diagram as STD
    ad = {"logpfc_hip": LOG, "stdpfc_hip": STD}[slug]
    rho = 0.63 if slug == "logpfc_hip" else -0.25
    params = dict(beta=0.5, g=-2.0) if slug == "logpfc_hip" else dict(beta=1.0, g=0.0)
    sim = ad.build(params, structure, REFS["cell"], H2L, -0.9 if slug == "logpfc_hip" else -0.25, rho, 1.0, 0.1)
    ad.run(sim, 300)
    d = ad.diagnose(sim, structure, REFS["cell"])
    h = ad.diagnose_state(ad.get_state(sim), structure, REFS["cell"])
    assert d.fitness == pytest.approx(h.fitness, abs=1e-13)
    assert d.amplitudes == pytest.approx(h.amplitudes, abs=1e-14)
    assert d.details == h.details


# -- adaptive dt batched; frozen-q solves batched across densities ------------------
def test_batched_adaptive_dt_equals_serial(hip_device):
    import time
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.phase_diagram import batched as B
    from pfc_hip.phase_diagram import q0star as Q
    params = {"beta": 0.5, "g": -2.0, "v0": 1.0}
    _, qscan = _settings()
    m = Q.MinimizeFreeEnergy(LOG, params, "triangle", SMALL, H2L, _frozen_relax(), qscan)
    pts = [(-0.9, 0.62, q) for q in (0.90, 0.95, 0.96, 1.0, 1.06)]
    t0 = time.perf_counter()
    serial = [m.evaluate(*p) for p in pts]
    t1 = time.perf_counter()
    batched = B.evaluate_batched(m, pts)
    t2 = time.perf_counter()
    print(f"\nadaptive dt, {len(pts)} points: serial {t1 - t0:.2f} s, batched {t2 - t1:.2f} s; "
          f"dt_max {[round(e.dt_max, 4) for e in batched]} steps {[e.steps for e in batched]}", flush=True)
    for s, b in zip(serial, batched):
        assert (s.f, s.fitness, s.amplitudes, s.stop, s.steps, s.dt_max) == \
            (b.f, b.fitness, b.amplitudes, b.stop, b.steps, b.dt_max)
        assert np.array_equal(s.state, b.state)
    assert max(e.dt_max for e in batched) > 0.01


def test_frozen_q_batched_across_densities(hip_device):
    import time
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.phase_diagram import frozen_q as FQ
    params = {"beta": 0.5, "g": -2.0, "v0": 1.0}
    m = FQ.MinimizeFrozenQ(LOG, params, "triangle", SMALL, H2L, _frozen_relax())
    rhos = [0.58, 0.60, 0.62, 0.64, 0.66, 0.68, 0.70]
    t0 = time.perf_counter()
    serial = [m.solve(-0.9, r) for r in rhos]
    t1 = time.perf_counter()
    batched = FQ.solve_batched(m, -0.9, rhos)
    t2 = time.perf_counter()
    print(f"\nfrozen-q {len(rhos)} densities: serial {t1 - t0:.2f} s, batched {t2 - t1:.2f} s; "
          f"statuses {[r['status'] for r in batched]}", flush=True)
    keys = ("status", "q", "q_relaxed", "f", "f_frozen", "fitness", "stop", "steps", "iters", "A1", "A2", "A3",
            "peak_spread", "subq_ratio", "n_peaks", "defect_flag")
    for a, b in zip(serial, batched):
        assert {k: a.get(k) for k in keys} == {k: b.get(k) for k in keys}
