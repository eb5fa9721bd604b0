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

    def solve(self, temp, meanDensity, q_start=1.0, state=None):
        """One density; returns a record (never raises: an exception is
        status 'error'): status, q (= q_opt of the final field, the q*),
        q_relaxed (the box it was relaxed in), f (at q_relaxed), f_frozen
        (frozen f at q), fitness, A1.., stop, steps, iters, wall_s, q_start,
        message, method, and the adapter's defect diagnostics if any."""
        t0 = time.perf_counter()
        st = self.settings
        last_stage = len(st.schedule) - 1
        q, carry, steps, stage = float(q_start), state, 0, 0
        rec = {'method': METHOD, 'q_start': float(q_start), 'message': ''}
        last, field = None, None
        try:
            for it in range(st.maxIter):
                m = self.stage_minimizer(stage)
                f, fit, *amps = m.Get_f_ABC(temp, meanDensity, q, state=carry)
                steps += m.last_steps
                field = m.last_state
                frozen = self.adapter.frozen_field(field, self.params, temp, self.cell)
                q_new = float(frozen.q_opt())
                last = {'q': q, 'q_opt': q_new, 'f': float(f), 'f_frozen': float(frozen.f(q_new)),
                        'fitness': float(fit), 'amps': [float(a) for a in amps], 'stop': m.last_stop, 'it': it}
                if m.last_stop not in (_q.STOP_CONVERGED, _q.STOP_TOTALT):
                    break
                dq = q_new - q
                carry, q = field, q_new
                if stage < last_stage:
                    stage += 1
                elif abs(dq) < st.qTol and m.last_stop == _q.STOP_CONVERGED:
                    break
        except Exception as e:  # noqa: BLE001 - a density's failure is its status
            rec.update(status=STATUS_ERROR, message=f'{type(e).__name__}: {e}', steps=steps,
                       wall_s=time.perf_counter() - t0)
            self.last_result, self.last_state = rec, None
            return rec
        settled = (last['stop'] == _q.STOP_CONVERGED and stage == last_stage
                   and abs(last['q_opt'] - last['q']) < st.qTol)
        if settled:
            status = STATUS_OK
        elif last['stop'] in (_q.STOP_CONVERGED, _q.STOP_TOTALT):
            status = STATUS_NOT_CONVERGED
        else:
            status = last['stop']
        rec.update(status=status, q=last['q_opt'], q_relaxed=last['q'], f=last['f'], f_frozen=last['f_frozen'],
                   fitness=last['fitness'], stop=last['stop'], steps=steps, iters=last['it'] + 1)
        for j, a in enumerate(last['amps']):
            rec[f'A{j + 1}'] = a
        defects = getattr(self.adapter, 'defect_diagnostics', None)
        if defects is not None and field is not None:
            rec.update(defects(field, self.cell))
        rec['wall_s'] = time.perf_counter() - t0
        self.last_result, self.last_state = rec, field
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
def _ok(r):
    return r is not None and r.get('status') == STATUS_OK


def edge_walks(grid, results):
    """(i, d) for every ok index i of the sorted density `grid` whose
    neighbour i + d (d = -1, +1) is not ok; `results` maps density -> record."""
    out = []
    for i, p in enumerate(grid):
        if _ok(results.get(p)):
            out += [(i, d) for d in (-1, 1) if 0 <= i + d < len(grid) and not _ok(results.get(grid[i + d]))]
    return out


def edge_walk(solve, grid, results, i, d, maxPast=3):
    """Walk from ok index i in direction d: each non-ok density is solved
    again by solve(phi, q_start) with q_start the q* of grid[i]; stop at an
    originally ok density, the grid's end or after maxPast consecutive
    non-ok results.  Returns the records, with q_start, edge_from, edge_dir."""
    q_seed = float(results[grid[i]]['q'])
    out, fails, j = [], 0, i + d
    while 0 <= j < len(grid) and fails < maxPast and not _ok(results.get(grid[j])):
        r = dict(solve(grid[j], q_seed))
        r.update(phi=grid[j], q_start=q_seed, edge_from=grid[i], edge_dir=d, pass_='edge')
        out.append(r)
        fails = 0 if _ok(r) else fails + 1
        j += d
    return out


def merge_edge(results, edge_records):
    """`results` with each non-ok original replaced by an ok edge record (the
    lowest f when several edge records are ok)."""
    out = dict(results)
    for r in edge_records:
        if not _ok(r):
            continue
        cur = out.get(r['phi'])
        if not _ok(cur) or (cur.get('pass_') == 'edge' and r['f'] < cur['f']):
            out[r['phi']] = r
    return out


def scan_density(minimizer, temp, densities, q_start=1.0, record=None, done=None):
    """One structure at one temperature over strictly increasing `densities`:
    independent solves (skipping those in `done`, density -> record from a
    checkpoint), then the edge pass if the settings ask for it.  `record`
    receives every record as it is made.  Returns (final, pass1, edge):
    density -> merged record, density -> independent record, edge records."""
    grid = [float(p) for p in densities]
    if any(b <= a for a, b in zip(grid, grid[1:])):
        raise ValueError('densities must be strictly increasing')
    pass1 = dict(done or {})
    for p in grid:
        if p in pass1:
            continue
        r = minimizer.solve(temp, p, q_start)
        r.update(phi=p, pass_='indep')
        pass1[p] = r
        if record is not None:
            record(r)
    edge = []
    if minimizer.settings.edge:
        def solve(p, q):
            r = minimizer.solve(temp, p, q)
            if record is not None:
                record({**r, 'phi': p, 'pass_': 'edge'})
            return r
        for i, d in edge_walks(grid, pass1):
            edge += edge_walk(solve, grid, pass1, i, d, minimizer.settings.maxPast)
    return merge_edge(pass1, edge), pass1, edge


# ---------------------------------------------------------------------------
# many densities at once
# ---------------------------------------------------------------------------
def solve_batched(minimizer, temp, densities, q_start=1.0, precision=64):
    """MinimizeFrozenQ.solve at every density, in lockstep: iteration `it`
    relaxes every density still iterating as one batch (evaluate_batched,
    fields carried, each at its own q and dt) at the stage every density is
    in at that iteration, then jumps each to the q_opt of its relaxed field.
    The records equal solve()'s (wall_s is the batch's elapsed time).  If a
    batch raises, the densities still iterating are solved one at a time."""
    from . import batched as _b
    t0 = time.perf_counter()
    st = minimizer.settings
    last_stage = len(st.schedule) - 1
    ad = minimizer.adapter
    rhos = [float(p) for p in densities]
    n = len(rhos)
    q = [float(q_start)] * n
    carry = [None] * n
    steps = [0] * n
    last = [None] * n
    field = [None] * n
    done = [False] * n
    stage = 0
    fallback = []
    for it in range(st.maxIter):
        run = [i for i in range(n) if not done[i]]
        if not run:
            break
        m = minimizer.stage_minimizer(stage)
        try:
            evs = _b.evaluate_batched(m, [(temp, rhos[i], q[i]) for i in run], states=[carry[i] for i in run],
                                      precision=precision)
        except Exception:  # noqa: BLE001 - those densities are solved one at a time below
            fallback = run
            break
        for i, ev in zip(run, evs):
            steps[i] += ev.steps
            field[i] = ev.state
            frozen = ad.frozen_field(ev.state, minimizer.params, temp, minimizer.cell)
            q_new = float(frozen.q_opt())
            # 'stage': the stage solve() holds after this iteration (it
            # advances before the next relaxation unless the loop stops here)
            last[i] = {'q': q[i], 'q_opt': q_new, 'f': float(ev.f), 'f_frozen': float(frozen.f(q_new)),
                       'fitness': float(ev.fitness), 'amps': [float(a) for a in ev.amplitudes], 'stop': ev.stop,
                       'it': it, 'stage': stage}
            if ev.stop not in (_q.STOP_CONVERGED, _q.STOP_TOTALT):
                done[i] = True
                continue
            dq = q_new - q[i]
            carry[i], q[i] = ev.state, q_new
            if stage < last_stage:
                last[i]['stage'] = stage + 1
            elif abs(dq) < st.qTol and ev.stop == _q.STOP_CONVERGED:
                done[i] = True
        if stage < last_stage:
            stage += 1
    out = []
    wall = time.perf_counter() - t0
    for i in range(n):
        if i in fallback:
            out.append(minimizer.solve(temp, rhos[i], q_start))
            continue
        L = last[i]
        rec = {'method': METHOD, 'q_start': float(q_start), 'message': ''}
        settled = L['stop'] == _q.STOP_CONVERGED and L['stage'] == last_stage and abs(L['q_opt'] - L['q']) < st.qTol
        if settled:
            status = STATUS_OK
        elif L['stop'] in (_q.STOP_CONVERGED, _q.STOP_TOTALT):
            status = STATUS_NOT_CONVERGED
        else:
            status = L['stop']
        rec.update(status=status, q=L['q_opt'], q_relaxed=L['q'], f=L['f'], f_frozen=L['f_frozen'],
                   fitness=L['fitness'], stop=L['stop'], steps=steps[i], iters=L['it'] + 1)
        for j, a in enumerate(L['amps']):
            rec[f'A{j + 1}'] = a
        defects = getattr(ad, 'defect_diagnostics', None)
        if defects is not None and field[i] is not None:
            rec.update(defects(field[i], minimizer.cell))
        rec['wall_s'] = wall
        out.append(rec)
    return out


def scan_density_batched(minimizer, temp, densities, q_start=1.0, record=None, precision=64):
    """scan_density with the independent solves batched (solve_batched); the
    edge walks, few and sequential by nature, run one solve at a time.
    Returns (final, pass1, edge) as scan_density does."""
    grid = [float(p) for p in densities]
    if any(b <= a for a, b in zip(grid, grid[1:])):
        raise ValueError('densities must be strictly increasing')
    pass1 = {}
    for p, r in zip(grid, solve_batched(minimizer, temp, grid, q_start, precision=precision)):
        r.update(phi=p, pass_='indep')
        pass1[p] = r
        if record is not None:
            record(r)
    edge = []
    if minimizer.settings.edge:
        def solve(p, q):
            r = minimizer.solve(temp, p, q)
            if record is not None:
                record({**r, 'phi': p, 'pass_': 'edge'})
            return r
        for i, d in edge_walks(grid, pass1):
            edge += edge_walk(solve, grid, pass1, i, d, minimizer.settings.maxPast)
    return merge_edge(pass1, edge), pass1, edge
