"""Frozen-field q0* minimisation, the alternative to Find_q0Star's q00 scan
(the reference's phase_diagram/frozen_q.py: MinimizeFrozenQ :68-166, the edge
pass :169-211, scan_density :214-243).

For a field held fixed on the grid only the gradient term of f depends on
the box scale q (every wavevector scales as q k_ref), so f(q) is a closed
form with a closed-form minimiser q_opt (the adapter's `frozen_field`).  At
a relaxed equilibrium q_opt = q (envelope theorem).  `solve` alternates:
relax at q (Get_f_ABC, rollbacks and stops included) to the stage's fTol,
jump q -> q_opt of the relaxed field (carried, the box rescaled), and move
to the next, tighter fTol of `schedule`; it ends when |q_opt - q| < qTol at
the last stage with a converged stop, after maxIter iterations, or when a
relaxation is rejected.

Statuses: 'ok'; the q0star stop names for a rejected relaxation ('poor_fit',
'min_amplitude', 'max_depth', 'nan'); 'not_converged' (q or f still moving
after maxIter); 'error' (an exception, message 'Type: text').

`scan_density` runs one density list: independent solves from q_start,
then the edge pass: every non-ok density next to an ok one is solved again
from a fresh ansatz started at the ok neighbour's q*, walking outward until
an ok original, the grid's end or maxPast consecutive failures; an ok edge
result replaces a non-ok original (the lowest f among several).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, replace

import numpy as np

from . import q0star as _q

METHOD = 'frozen-q'
STATUS_OK = 'ok'
STATUS_NOT_CONVERGED = 'not_converged'
STATUS_ERROR = 'error'
DEFAULT_SCHEDULE = (1e-6, 1e-8, 1e-10, 1e-12)


class FrozenQRejected(Exception):
    """The solve gave no accepted state; `status` holds its code."""

    def __init__(self, status, message=''):
        super().__init__(message or status)
        self.status = status


@dataclass(frozen=True)
class FrozenQSettings:
    """schedule: fTol per stage, the last the final one (relax.fTol is not
    used); qTol: |q_opt - q| at the final stage for convergence; maxIter:
    relax/jump iterations; edge, maxPast: scan_density's edge pass."""
    schedule: tuple = DEFAULT_SCHEDULE
    qTol: float = 1e-8
    maxIter: int = 40
    edge: bool = True
    maxPast: int = 3


class _FrozenRun:
    """The relax/jump iteration of one density as a state machine, driven one
    relaxation at a time by the serial solve or in lockstep with other
    densities by solve_batched.  `box` is the scale the next relaxation runs
    at, `carry` the field it starts from (None: a fresh ansatz), `stage` its
    index into the fTol schedule."""

    def __init__(self, settings, q_start, state=None):
        self.settings = settings
        self.q_start = float(q_start)
        self.box = float(q_start)
        self.carry = state
        self.stage = 0
        self.steps = 0
        self.seen = None       # the last relaxation: box, q_opt, f, f_frozen, fitness, amps, stop, iteration
        self.field = None
        self.finished = False
        self.error = None

    @property
    def final_stage(self):
        return len(self.settings.schedule) - 1

    def observe(self, it, f, fitness, amps, stop, steps, field, frozen):
        """Take one relaxation's outcome; decide the next box and stage."""
        self.steps += steps
        self.field = field
        q_opt = float(frozen.q_opt())
        self.seen = dict(box=self.box, q_opt=q_opt, f=float(f), f_frozen=float(frozen.f(q_opt)),
                         fitness=float(fitness), amps=[float(a) for a in amps], stop=stop, it=it)
        if stop not in (_q.STOP_CONVERGED, _q.STOP_TOTALT):
            self.finished = True           # a rejected relaxation ends the density
            return
        jump = q_opt - self.box
        self.carry, self.box = field, q_opt  # the field follows the box to q_opt
        if self.stage < self.final_stage:
            self.stage += 1
        elif stop == _q.STOP_CONVERGED and abs(jump) < self.settings.qTol:
            self.finished = True

    def record(self, adapter, cell, wall_s):
        rec = {'method': METHOD, 'q_start': self.q_start, 'message': ''}
        if self.error is not None:
            rec.update(status=STATUS_ERROR, message=self.error, steps=self.steps, wall_s=wall_s)
            return rec
        s = self.seen
        # stage: where the run stands after its last relaxation; accepted
        # only at the final stage, converged and with q_opt within qTol of
        # the box the field was relaxed in
        if s['stop'] == _q.STOP_CONVERGED and self.stage == self.final_stage and \
                abs(s['q_opt'] - s['box']) < self.settings.qTol:
            status = STATUS_OK
        elif s['stop'] in (_q.STOP_CONVERGED, _q.STOP_TOTALT):
            status = STATUS_NOT_CONVERGED
        else:
            status = s['stop']
        rec.update(status=status, q=s['q_opt'], q_relaxed=s['box'], f=s['f'], f_frozen=s['f_frozen'],
                   fitness=s['fitness'], stop=s['stop'], steps=self.steps, iters=s['it'] + 1)
        rec.update({f'A{j + 1}': a for j, a in enumerate(s['amps'])})
        defects = getattr(adapter, 'defect_diagnostics', None)
        if defects is not None and self.field is not None:
            rec.update(defects(self.field, cell))
        rec['wall_s'] = wall_s
        return rec


class MinimizeFrozenQ:
    """MinimizeFreeEnergy's constructor arguments with FrozenQSettings in
    place of the q-scan settings.  The adapter must provide
    `frozen_field(state, params, temp, cell)` (an object with f(q) and
    q_opt()); `defect_diagnostics(state, cell)` is recorded when present."""

    def __init__(self, adapter, params, structure, cell, ansatz, relax, settings=None):
        if not hasattr(adapter, 'frozen_field'):
            raise ValueError(f'adapter {getattr(adapter, "SLUG", adapter)!r} has no frozen_field')
        self.adapter = adapter
        self.params = adapter.validate_params(params)
        self.structure = structure
        self.cell = adapter.validate_cell(cell)
        self.ansatz = adapter.validate_ansatz(ansatz)
        self.relax = relax
        self.settings = settings or FrozenQSettings()
        if len(self.settings.schedule) < 1:
            raise ValueError('schedule needs at least one fTol')
        self.last_result = None
        self.last_state = None

    def stage_minimizer(self, stage):
        """The relaxation of one schedule stage: its fTol, and ATol loosened
        to 1e6 fTol before the final stage."""
        sched = self.settings.schedule
        fTol = float(sched[stage])
        final = stage == len(sched) - 1
        relax = replace(self.relax, fTol=fTol, ATol=self.relax.ATol if final else max(self.relax.ATol, 1e6 * fTol))
        return _q.MinimizeFreeEnergy(self.adapter, self.params, self.structure, self.cell, self.ansatz, relax,
                                     _q.QScanSettings(1, 0.0, 1, 1.0))

    def _frozen(self, field, temp):
        return self.adapter.frozen_field(field, self.params, temp, self.cell)

    def solve(self, temp, meanDensity, q_start=1.0, state=None):
        """One density; returns a record (never raises: an exception is
        status 'error'): status, q (= q_opt of the final field, the q*),
        q_relaxed (the box it was relaxed in), f (at q_relaxed), f_frozen
        (frozen f at q), fitness, A1.., stop, steps, iters, wall_s, q_start,
        message, method, and the adapter's defect diagnostics if any."""
        t0 = time.perf_counter()
        run = _FrozenRun(self.settings, q_start, state)
        try:
            it = 0
            while it < self.settings.maxIter and not run.finished:
                m = self.stage_minimizer(run.stage)
                f, fit, *amps = m.Get_f_ABC(temp, meanDensity, run.box, state=run.carry)
                run.observe(it, f, fit, amps, m.last_stop, m.last_steps, m.last_state,
                            self._frozen(m.last_state, temp))
                it += 1
        except Exception as e:  # noqa: BLE001 - a density's failure is its status
            run.error = f'{type(e).__name__}: {e}'
        rec = run.record(self.adapter, self.cell, time.perf_counter() - t0)
        self.last_result, self.last_state = rec, (None if run.error else run.field)
        return rec

    def Find_q0Star(self, temp, meanDensity, q_start=1.0, record=None):
        """(q0*, fmin) as MinimizeFreeEnergy.Find_q0Star gives them; raises
        FrozenQRejected unless the state is accepted.  The whole record is
        in `last_result` and goes to `record` when given."""
        r = self.solve(temp, meanDensity, q_start)
        if record is not None:
            record(r)
        if r['status'] != STATUS_OK:
            raise FrozenQRejected(r['status'], r.get('message', ''))
        return r['q'], r['f']


# ---------------------------------------------------------------------------
# the edge pass
# ---------------------------------------------------------------------------
def _accepted(rec):
    return rec is not None and rec.get('status') == STATUS_OK


def _ok_mask(grid, results):
    return np.array([_accepted(results.get(p)) for p in grid], dtype=bool)


def edge_walks(grid, results):
    """(i, d) for every ok index i of the sorted density `grid` whose
    neighbour i + d (d = -1, +1) is not ok; `results` maps density -> record."""
    ok = _ok_mask(grid, results)
    n = len(ok)
    left = {int(i) for i in np.flatnonzero(ok[1:] & ~ok[:-1]) + 1} if n > 1 else set()    # i ok, i-1 not
    right = {int(i) for i in np.flatnonzero(ok[:-1] & ~ok[1:])} if n > 1 else set()       # i ok, i+1 not
    return sorted([(i, -1) for i in left] + [(i, 1) for i in right])


def edge_walk(solve, grid, results, i, d, maxPast=3):
    """Walk from ok index i in direction d: each non-ok density is solved
    again by solve(phi, q_start) with q_start the q* of grid[i]; stop at an
    originally ok density, the grid's end or after maxPast consecutive
    non-ok results.  Returns the records, with q_start, edge_from, edge_dir."""
    seed = float(results[grid[i]]['q'])
    stop = len(grid) if d > 0 else -1
    out, misses = [], 0
    for j in range(i + d, stop, d):
        if misses >= maxPast or _accepted(results.get(grid[j])):
            break
        rec = {**solve(grid[j], seed), 'phi': grid[j], 'q_start': seed, 'edge_from': grid[i], 'edge_dir': d,
               'pass_': 'edge'}
        out.append(rec)
        misses = 0 if _accepted(rec) else misses + 1
    return out


def merge_edge(results, edge_records):
    """`results` with each non-ok original replaced by an ok edge record (the
    lowest f when several edge records are ok, the first on a tie)."""
    best = {}
    for r in edge_records:
        if _accepted(r) and (r['phi'] not in best or r['f'] < best[r['phi']]['f']):
            best[r['phi']] = r
    out = dict(results)
    out.update({p: r for p, r in best.items() if not _accepted(results.get(p))})
    return out


def _edge_pass(minimizer, temp, grid, pass1, record):
    """The edge walks of scan_density (one solve at a time)."""
    if not minimizer.settings.edge:
        return []

    def solve(p, q):
        r = minimizer.solve(temp, p, q)
        if record is not None:
            record({**r, 'phi': p, 'pass_': 'edge'})
        return r

    edge = []
    for i, d in edge_walks(grid, pass1):
        edge += edge_walk(solve, grid, pass1, i, d, minimizer.settings.maxPast)
    return edge


def _strict_grid(densities):
    grid = [float(p) for p in densities]
    if any(b <= a for a, b in zip(grid, grid[1:])):
        raise ValueError('densities must be strictly increasing')
    return grid


def scan_density(minimizer, temp, densities, q_start=1.0, record=None, done=None):
    """One structure at one temperature over strictly increasing `densities`:
    independent solves (skipping those in `done`, density -> record from a
    checkpoint), then the edge pass if the settings ask for it.  `record`
    receives every record as it is made.  Returns (final, pass1, edge):
    density -> merged record, density -> independent record, edge records."""
    grid = _strict_grid(densities)
    pass1 = dict(done or {})
    for p in (p for p in grid if p not in pass1):
        pass1[p] = {**minimizer.solve(temp, p, q_start), 'phi': p, 'pass_': 'indep'}
        if record is not None:
            record(pass1[p])
    edge = _edge_pass(minimizer, temp, grid, pass1, record)
    return merge_edge(pass1, edge), pass1, edge


# ---------------------------------------------------------------------------
# many densities at once
# ---------------------------------------------------------------------------
def solve_batched(minimizer, temp, densities, q_start=1.0, precision=64):
    """MinimizeFrozenQ.solve at every density, in lockstep: iteration `it`
    relaxes every density still iterating as one batch (evaluate_batched,
    fields carried, each at its own q and dt), then each density takes its
    outcome as solve() does (the same _FrozenRun).  Every density is at the
    same stage at the same iteration, so the batch shares one stage's
    settings.  The records equal solve()'s (wall_s is the batch's elapsed
    time).  If a batch raises, the densities still iterating are solved one
    at a time."""
    from . import batched as _b
    t0 = time.perf_counter()
    rhos = [float(p) for p in densities]
    runs = [_FrozenRun(minimizer.settings, q_start) for _ in rhos]
    fallback = []
    for it in range(minimizer.settings.maxIter):
        live = [k for k, r in enumerate(runs) if not r.finished]
        if not live:
            break
        m = minimizer.stage_minimizer(runs[live[0]].stage)
        try:
            evs = _b.evaluate_batched(m, [(temp, rhos[k], runs[k].box) for k in live],
                                      states=[runs[k].carry for k in live], precision=precision)
        except Exception:  # noqa: BLE001 - those densities are solved one at a time below
            fallback = live
            break
        for k, ev in zip(live, evs):
            runs[k].observe(it, ev.f, ev.fitness, ev.amplitudes, ev.stop, ev.steps, ev.state,
                            minimizer._frozen(ev.state, temp))
    wall = time.perf_counter() - t0
    return [minimizer.solve(temp, rhos[k], q_start) if k in fallback
            else runs[k].record(minimizer.adapter, minimizer.cell, wall) for k in range(len(rhos))]


def scan_density_batched(minimizer, temp, densities, q_start=1.0, record=None, precision=64):
    """scan_density with the independent solves batched (solve_batched); the
    edge walks, few and sequential by nature, run one solve at a time.
    Returns (final, pass1, edge) as scan_density does."""
    grid = _strict_grid(densities)
    pass1 = {}
    for p, r in zip(grid, solve_batched(minimizer, temp, grid, q_start, precision=precision)):
        pass1[p] = {**r, 'phi': p, 'pass_': 'indep'}
        if record is not None:
            record(pass1[p])
    edge = _edge_pass(minimizer, temp, grid, pass1, record)
    return merge_edge(pass1, edge), pass1, edge
