"""frozen_q (pfc_hip.phase_diagram.frozen_q) on the CPU: the frozen-field
f(q) of the hip adapters against the oracle's free energy of the same field
in the box of each q, its minimiser against brute force, the edge-pass
helpers, and the relax/jump loop on a toy adapter.  The reference pins the
same behaviours in tests/test_phase_diagram_frozen_q.py:22-83; the
evaluations on the hip path are in test_gpu_phase_diagram.py."""
import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip.models import _adapter as A
from pfc_hip.models.logpfc_hip import phase_diagram as LOG
from pfc_hip.models.stdpfc_hip import phase_diagram as STD
from pfc_hip.phase_diagram import frozen_q as FQ
from pfc_hip.phase_diagram import q0star as Q

CELL = {"uNx": 14, "uNy": 24, "Mx": 2, "My": 1}
H2L = {"mode": "H2L", "A1": 0.9, "A2": 0.0, "A3": 0.0}


def _oracle_fe(model, phi, params, temp, q):
    g = A.box(CELL, q)
    geom = O.build_geometry(g.Nx, g.Ny, g.Lx, g.Ly)
    sim = O.Simulation(O.Params(model=model, beta=params["beta"], q0=1.0, epsilon=temp, g=params["g"]), geom)
    sim.set_field(phi)
    return sim.free_energy()


@pytest.mark.parametrize("beta,g,temp", [(0.5, -2.0, -0.9), (0.25, 1.0, -3.0), (1.0, 0.0, -2.0)])
def test_log_frozen_f_is_the_free_energy_in_each_box(beta, g, temp):
    params = {"beta": beta, "g": g, "v0": 1.0}
    rng = np.random.default_rng(1)
    base = LOG.initial_field("triangle", CELL, H2L, 0.7)
    phi = base * (1 + 0.05 * rng.standard_normal(base.shape))
    ff = LOG.frozen_field(phi, params, temp, CELL)
    for q in (0.7, 0.93, 1.0, 1.08, 1.3):
        fo = _oracle_fe("log", phi, params, temp, q)
        assert abs(float(ff.f(q)) - fo) <= 1e-12 * max(1.0, abs(fo)), q
    qs = np.linspace(0.05, 2.0, 39001)
    assert abs(qs[np.argmin(ff.f(qs))] - ff.q_opt()) < 1e-4
    # a one-mode ansatz sits on |k| = 1: q_opt = 1
    assert abs(LOG.frozen_field(base, params, temp, CELL).q_opt() - 1.0) < 1e-12


def test_std_frozen_f_is_the_free_energy_in_each_box():
    params = {"beta": 1.0, "g": 0.5, "v0": 1.0}
    rng = np.random.default_rng(2)
    n = STD.initial_field("triangle", CELL, {"mode": "x", "A1": 0.3, "A2": 0.05, "A3": 0.01}, -0.2)
    n = n + 0.02 * rng.standard_normal(n.shape)
    ff = STD.frozen_field(n, params, -0.3, CELL)
    for q in (0.8, 1.0, 1.2):
        fo = _oracle_fe("std", n, params, -0.3, q)
        assert abs(float(ff.f(q)) - fo) <= 1e-13 * max(1.0, abs(fo)), q
    # q_opt is the stationary point of the quartic in q
    h = 1e-5
    q = ff.q_opt()
    assert abs(float(ff.f(q + h) - ff.f(q - h)) / (2 * h)) < 1e-8


def test_defect_diagnostics_on_clean_and_defective_fields():
    clean = LOG.initial_field("triangle", CELL, H2L, 0.7)
    d = LOG.defect_diagnostics(clean, CELL)
    assert d["n_peaks"] == 4 and d["peak_spread"] < 1e-12 and d["subq_ratio"] < 1e-20 and not d["defect_flag"]
    ny, nx = clean.shape
    x = np.arange(nx)[None, :] * np.ones((ny, 1))
    dirty = clean + 0.05 * np.cos(2 * np.pi * x / nx)        # a long-wave mode, |k_ref| < 0.95
    d = LOG.defect_diagnostics(dirty, CELL)
    assert d["subq_ratio"] > 1e-3 and d["defect_flag"]


def test_edge_helpers():
    grid = [0.1, 0.2, 0.3, 0.4, 0.5]
    res = {0.1: {"status": "poor_fit"}, 0.2: {"status": "ok", "q": 0.97, "f": -1.0},
           0.3: {"status": "ok", "q": 0.98, "f": -1.1}, 0.4: {"status": "poor_fit"}, 0.5: {"status": "poor_fit"}}
    assert FQ.edge_walks(grid, res) == [(1, -1), (2, 1)]
    calls = []

    def solve(p, q):
        calls.append((p, q))
        return {"status": "ok" if p == 0.4 else "poor_fit", "q": q, "f": -2.0}
    recs = FQ.edge_walk(solve, grid, res, 2, 1, maxPast=3)
    assert calls == [(0.4, 0.98), (0.5, 0.98)]
    merged = FQ.merge_edge(res, recs)
    assert merged[0.4]["status"] == "ok" and merged[0.4]["edge_from"] == 0.3
    assert merged[0.5]["status"] == "poor_fit"
    # maxPast consecutive failures end a walk
    calls.clear()
    FQ.edge_walk(lambda p, q: calls.append(p) or {"status": "poor_fit"}, grid, res, 2, 1, maxPast=1)
    assert calls == [0.4]


# -- the relax/jump loop on a toy: the relaxed field's q_opt moves toward a
#    fixed lattice scale QF; every relaxation converges in one block
QF = 1.037


class _Sim:
    def __init__(self, q, state):
        self.q, self.t, self.dt = q, 0.0, 0.1
        self.x = np.array(state if state is not None else [1.0, 0.0], dtype=float)


class ToyFrozen:
    SLUG = "toyfrozen"
    STRUCTURES = {"dots": "dots"}

    def __init__(self, fitness=1.0, fail_at=None):
        self.fitness, self.fail_at = fitness, fail_at
        self.qs = []

    def validate_params(self, p):
        return dict(p)

    validate_cell = validate_ansatz = validate_params

    def build(self, params, structure, cell, ansatz, temp, meanDensity, q00, dt, state=None, nThreads=1):
        if self.fail_at is not None and meanDensity == self.fail_at:
            raise ValueError("no such state")
        self.qs.append(q00)
        return _Sim(q00, state)

    def get_state(self, s):
        return s.x.copy()

    def set_state(self, s, v):
        s.x[...] = v

    def get_time(self, s):
        return s.t

    def set_time(self, s, t):
        s.t = t

    def get_dt(self, s):
        return s.dt

    def step(self, s):
        s.t += s.dt

    def free_energy(self, s):
        return (s.q - QF) ** 2 - 1.0

    def diagnose(self, s, structure, cell):
        return Q.Diagnostics(fitness=self.fitness, amplitudes=(0.3,))

    def frozen_field(self, state, params, temp, cell):
        # the relaxed field at scale q prefers q_opt = q + (QF - q) / 2
        q = self.qs[-1]

        class F:
            def q_opt(self_):
                return q + (QF - q) / 2

            def f(self_, x):
                return (x - QF) ** 2 - 1.0
        return F()


RELAX = Q.RelaxSettings(dt=0.1, totalT=1e3, stepT=10.0, divs=10, maxDepth=2, requiredFitness=0.99, minimumA=1e-6)


def test_solve_converges_on_the_fixed_point_through_the_schedule():
    m = FQ.MinimizeFrozenQ(ToyFrozen(), {}, "dots", {}, {}, RELAX, FQ.FrozenQSettings(qTol=1e-8, maxIter=60))
    r = m.solve(0.0, 0.5, q_start=1.0)
    assert r["status"] == FQ.STATUS_OK and abs(r["q"] - QF) < 2e-8
    assert r["iters"] >= len(FQ.DEFAULT_SCHEDULE) and r["stop"] == Q.STOP_CONVERGED
    assert r["f"] == pytest.approx(-1.0, abs=1e-15) and r["method"] == FQ.METHOD
    q, f = m.Find_q0Star(0.0, 0.5)
    assert abs(q - QF) < 2e-8


def test_solve_statuses():
    few = FQ.FrozenQSettings(maxIter=3)
    assert FQ.MinimizeFrozenQ(ToyFrozen(), {}, "dots", {}, {}, RELAX, few).solve(0.0, 0.5)["status"] == \
        FQ.STATUS_NOT_CONVERGED
    bad = FQ.MinimizeFrozenQ(ToyFrozen(fitness=0.5), {}, "dots", {}, {}, RELAX)
    assert bad.solve(0.0, 0.5)["status"] == Q.STOP_POOR_FIT
    with pytest.raises(FQ.FrozenQRejected):
        bad.Find_q0Star(0.0, 0.5)
    err = FQ.MinimizeFrozenQ(ToyFrozen(fail_at=0.5), {}, "dots", {}, {}, RELAX).solve(0.0, 0.5)
    assert err["status"] == FQ.STATUS_ERROR and err["message"] == "ValueError: no such state"
    with pytest.raises(ValueError):
        FQ.MinimizeFrozenQ(object(), {}, "dots", {}, {}, RELAX)


def test_stage_tolerances():
    m = FQ.MinimizeFrozenQ(ToyFrozen(), {}, "dots", {}, {}, RELAX)
    first, last = m.stage_minimizer(0).relax, m.stage_minimizer(len(FQ.DEFAULT_SCHEDULE) - 1).relax
    assert (first.fTol, first.ATol) == (1e-6, 1.0) and (last.fTol, last.ATol) == (1e-12, RELAX.ATol)


def test_scan_density_edge_pass_recovers_a_failed_density():
    class Picky(ToyFrozen):
        # density 0.3 fails from q_start = 1.0 but not when started near QF
        def build(self, params, structure, cell, ansatz, temp, meanDensity, q00, dt, state=None, nThreads=1):
            if meanDensity == 0.3 and state is None and q00 == 1.0:
                raise ValueError("no such state")
            return super().build(params, structure, cell, ansatz, temp, meanDensity, q00, dt, state, nThreads)
    m = FQ.MinimizeFrozenQ(Picky(), {}, "dots", {}, {}, RELAX, FQ.FrozenQSettings(maxIter=60))
    seen = []
    final, pass1, edge = FQ.scan_density(m, 0.0, [0.1, 0.2, 0.3], record=seen.append)
    assert pass1[0.3]["status"] == FQ.STATUS_ERROR and final[0.3]["status"] == FQ.STATUS_OK
    assert final[0.3]["pass_"] == "edge" and final[0.3]["edge_from"] == 0.2
    assert len(seen) == 4 and len(edge) == 1
    with pytest.raises(ValueError):
        FQ.scan_density(m, 0.0, [0.2, 0.1])


def test_solve_batched_equals_solve_on_a_batch_toy():
    # the batched driver's toy from test_phase_diagram_batched, given a
    # frozen field whose q_opt pulls toward QF: records equal solve()'s
    from test_phase_diagram_batched import Toy as BToy

    class FrozenToy(BToy):
        def frozen_field(self, state, params, temp, cell):
            x = float(np.mean(state[0]))

            class F_:
                def q_opt(self_):
                    return 1.037 + 0.01 * x

                def f(self_, qq):
                    return (qq - 1.037) ** 2 + x
            return F_()
    relax = Q.RelaxSettings(dt=0.1, totalT=2e3, stepT=10.0, divs=10, maxDepth=3, requiredFitness=0.995,
                            minimumA=1e-6, dtMax=0.5, dtRate=1e-3)
    m = FQ.MinimizeFrozenQ(FrozenToy(), {}, "dots", {}, {}, relax, FQ.FrozenQSettings(maxIter=12))
    rhos = [0.45, 0.5, 0.7, 2.5, 0.8]
    serial = [m.solve(0.2, r, 1.0) for r in rhos]
    batched = FQ.solve_batched(m, 0.2, rhos, 1.0)
    keys = ("status", "q", "q_relaxed", "f", "f_frozen", "fitness", "stop", "steps", "iters", "A1", "A2")
    for a, b in zip(serial, batched):
        assert {k: a.get(k) for k in keys} == {k: b.get(k) for k in keys}
    assert {r["status"] for r in batched} >= {FQ.STATUS_OK, Q.STOP_POOR_FIT}


def test_solve_batched_when_max_iter_ends_before_the_last_stage():
    from test_phase_diagram_batched import Toy as BToy

    class FrozenToy(BToy):
        def frozen_field(self, state, params, temp, cell):
            class F_:
                def q_opt(self_):
                    return 1.0

                def f(self_, qq):
                    return 0.0
            return F_()
    relax = Q.RelaxSettings(dt=0.1, totalT=2e3, stepT=10.0, divs=10, maxDepth=3, requiredFitness=0.995,
                            minimumA=1e-6)
    for iters in (2, 3, 4, 6):
        m = FQ.MinimizeFrozenQ(FrozenToy(), {}, "dots", {}, {}, relax, FQ.FrozenQSettings(maxIter=iters))
        a = [m.solve(0.2, r, 1.0) for r in (0.5, 0.45)]
        b = FQ.solve_batched(m, 0.2, [0.5, 0.45], 1.0)
        assert [x["status"] for x in a] == [x["status"] for x in b], iters
        assert [x["iters"] for x in a] == [x["iters"] for x in b], iters
