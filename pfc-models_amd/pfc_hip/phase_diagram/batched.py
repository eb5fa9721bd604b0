"""The phase-diagram driver's evaluations run as one device batch.

`evaluate_batched(minimizer, points)` performs Get_f_ABC (q0star.py) for
many (temp, meanDensity, q00) points at once: every point is a slot of one
device batch (its own epsilon, box and field; pfc_batch_set_slot_box), and
the slots advance together, one launch per kernel for all of them, block by
block.  The per-point semantics are the serial driver's
(MinimizeFreeEnergy.Get_f_ABC, the reference's q0star.py:329-406):

* the same fresh ansatz (read back from the device, then the seeded noise),
  the same dt = stepT / ceil(stepT / dt) and blocks of stepT;
* a block whose f rises by more than fTol |f| or turns NaN is handed, from
  the block's starting state, to the serial rollback (MinimizeFreeEnergy.
  _rollback: dt / divs sub-blocks, recursively to maxDepth) on a Simulation
  of its own, and the refined state goes back into the slot;
* the same stop tests after every block, with the same diagnostics on the
  same fields; a stopped slot leaves the active list.

With adaptive dt (relax.dtMax) each block is `dtChecks` sub-blocks and
every point keeps its own dt: the sub-blocks run in lockstep, each slot
taking its own number of steps in the same launch (pfc_batch_step_counts),
and a slot whose dt changes is rebuilt from its field as _adaptDt rebuilds
the Simulation (q0star.py:309-326).  Batches whose grids are too large for
the resident stepper cannot step slots by different counts; their points
are then evaluated one at a time.

`find_q0star_batched` runs one density's q00 scan with a whole pass per
batch; `scan_density_batched` runs the passes of a whole density list
together (pass p of every density still scanning in one batch), with
scan_density's results, statuses, records and checkpoint files.

The adapter provides, besides the serial names: `open_batch(params, cell,
capacity)`, an object with `place(slot, temp, meanDensity, q00, dt, state,
t=0.0)` (t: the simulation time the slot's model sees, if any),
`set_state(slot, state)`, `set_active(slots)`, `step(n)`, `step_each(counts)`
(False if the slots cannot take different step counts), and, for the
active slots in active order, `states()`, `free_energy()` and
`diagnose_active(structure, cell)` (the adapter's `diagnose` of each), and
`close()`; and `ansatz_state(structure, cell, ansatz, meanDensity)`, the
fresh ansatz `build` starts from.  The hip adapters read all fields with one
device call and compute the diagnostics on the device.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import q0star as Q
from ..models._base import advance_time


class _Point:
    """One evaluation's running values: time, f, A and diagnostics after the
    last block, time steps, stop reason (None while running), field, and
    with adaptive dt its dt, largest dt and sub-block A."""
    __slots__ = ("temp", "rho", "q00", "t", "f", "A", "diag", "steps", "stop", "state", "dt", "dt_max", "A_s")

    def __init__(self, temp, rho, q00):
        self.temp, self.rho, self.q00 = float(temp), float(rho), float(q00)
        self.t, self.steps, self.stop = 0.0, 0, None


def evaluate_batched(minimizer, points, states=None, precision=64):
    """Get_f_ABC at each (temp, meanDensity, q00) of `points` (optionally
    from `states` instead of the ansatz); returns one q0star.Evaluation per
    point, final state included."""
    m, ad, r = minimizer, minimizer.adapter, minimizer.relax
    points = [tuple(p) for p in points]
    if states is None:
        states = [None] * len(points)
    if not points:
        return []
    dt0 = r.stepT / int(np.ceil(r.stepT / r.dt))
    n = int(round(r.stepT / dt0))
    blocks = int(r.totalT / r.stepT)
    sb = ad.open_batch(m.params, m.cell, len(points), precision=precision)
    try:
        pts = [_Point(*p) for p in points]
        everyone = list(range(len(pts)))
        for j, (p, s) in enumerate(zip(pts, states)):
            if s is None:
                sb.place(j, p.temp, p.rho, p.q00, dt0, ad.ansatz_state(m.structure, m.cell, m.ansatz, p.rho))
            else:
                sb.place(j, p.temp, p.rho, p.q00, dt0, s)
        sb.set_active(everyone)
        if r.noise:  # the noise multiplies the ansatz as read back from the device, as _generateSim's does
            for j, x in enumerate(sb.states()):
                if states[j] is None:
                    sb.set_state(j, Q.noisy_state(x, r.seed, r.noise))
        for p, x, f, d in zip(pts, sb.states(), sb.free_energy(), sb.diagnose_active(m.structure, m.cell)):
            p.state, p.f, p.diag, p.A = x, float(f), d, d.amplitudes[0]
            p.dt = p.dt_max = dt0
        if r.dtMax is not None:
            if not _adaptive_blocks(m, ad, r, sb, pts, blocks):
                return [m.evaluate(t, rho, q, state=s) for (t, rho, q), s in zip(points, states)]
            return [Q.Evaluation(p.f, float(p.diag.fitness), tuple(float(a) for a in p.diag.amplitudes),
                                 p.stop or Q.STOP_TOTALT, p.steps, p.dt_max, p.state) for p in pts]
        for _ in range(blocks):
            act = [j for j, p in enumerate(pts) if p.stop is None]
            if not act:
                break
            sb.set_active(act)
            sb.step(n)
            fe = sb.free_energy()
            fields = sb.states()
            df = {}
            for k, j in enumerate(act):
                df[j] = _block(m, ad, r, sb, pts[j], j, float(fe[k]), fields[k], n, dt0, r.stepT)
            for j, d in zip(act, sb.diagnose_active(m.structure, m.cell)):
                p = pts[j]
                p.diag = d
                if p.stop is None:
                    _stop_tests(r, p, df[j], d)
        return [Q.Evaluation(p.f, float(p.diag.fitness), tuple(float(a) for a in p.diag.amplitudes),
                             p.stop or Q.STOP_TOTALT, p.steps, dt0, p.state) for p in pts]
    finally:
        sb.close()


def _block(m, ad, r, sb, p, j, f, field, n, dt, stepT):
    """Slot j after a block (of stepT) of n steps ending at free energy f and `field`:
    the serial driver's time and step bookkeeping, and its rollback of a
    rising or NaN block (on a Simulation of its own, from the block's
    starting state; the outcome goes back into the slot).  Returns the
    block's change of f."""
    t_before, f_before, before = p.t, p.f, p.state
    p.t = advance_time(p.t, dt, n)  # step by step, to the bit, as the serial run does
    p.steps += n
    df = f - f_before
    if not (np.isnan(f) or df > r.fTol * abs(f_before)):
        p.state, p.f = field, f
        return df
    sim = m._generateSim(p.temp, p.rho, p.q00, dt, before)
    ad.set_time(sim, t_before)
    steps0 = m.last_steps
    try:
        df = m._rollback(sim, p.temp, p.rho, p.q00, stepT, before, t_before, f_before, bool(np.isnan(f)),
                         divs=r.divs)
    except Q.MaxDepthReached as e:
        p.stop = Q.STOP_NAN if e.nan else Q.STOP_MAX_DEPTH
    finally:
        p.steps += m.last_steps - steps0
    p.state, p.f, p.t = ad.get_state(sim), float(ad.free_energy(sim)), ad.get_time(sim)
    sb.set_state(j, p.state)
    return df


def _adaptive_blocks(m, ad, r, sb, pts, blocks):
    """Get_f_ABC's adaptive-dt blocks for every point (q0star.py:347-404):
    each block is dtChecks sub-blocks of stepT/dtChecks, each slot at its own
    dt, dt adapted after every sub-block but the last and between blocks.
    Returns False, having stepped nothing, if the batch cannot step its slots
    by different counts."""
    sub_T = r.stepT / r.dtChecks
    for _ in range(blocks):
        act = [j for j, p in enumerate(pts) if p.stop is None]
        if not act:
            break
        f0 = {j: pts[j].f for j in act}
        for j in act:
            pts[j].A_s = pts[j].A
        for k in range(r.dtChecks):
            live = [j for j in act if pts[j].stop is None]
            if not live:
                break
            sb.set_active(live)
            counts = [int(round(sub_T / pts[j].dt)) for j in live]
            if not sb.step_each(counts):
                return False
            fe, fields = sb.free_energy(), sb.states()
            d_sub = {j: _block(m, ad, r, sb, pts[j], j, float(fe[i]), fields[i], counts[i], pts[j].dt, sub_T)
                     for i, j in enumerate(live)}
            for j, d in zip(live, sb.diagnose_active(m.structure, m.cell)):
                p = pts[j]
                p.diag = d
                if p.stop is not None:
                    continue
                A_new = d.amplitudes[0]
                if k < r.dtChecks - 1:
                    _adapt(m, r, sb, p, j, sub_T, d_sub[j], p.f, A_new, p.A_s)
                p.A_s = A_new
        for j in act:
            p = pts[j]
            if p.stop is not None:
                continue
            df = p.f - f0[j]
            A = p.diag.amplitudes[0]
            A_last = p.A
            _stop_tests(r, p, df, p.diag)
            if p.stop is None:
                _adapt(m, r, sb, p, j, sub_T, df / r.dtChecks, p.f, A, A_last)
    return True


def _adapt(m, r, sb, p, j, sub_T, delta_f, f, A, A_last):
    """_adaptDt for slot j: the next dt from the relative rate; on a change
    the slot is rebuilt at the new dt from its field, its time kept."""
    rate = max(abs(delta_f) / max(abs(f), 1e-300), abs(A - A_last) / max(A, 1e-300)) / sub_T
    target = r.dtMax / np.sqrt(1.0 + (rate / r.dtRate) ** 2)
    target = min(max(target, r.dt), r.dtMax, p.dt * r.dtGrow)
    new = sub_T / int(np.ceil(sub_T / target - 1e-9))
    if abs(new - p.dt) <= 1e-12 * p.dt:
        return
    sb.place(j, p.temp, p.rho, p.q00, new, p.state, t=p.t)
    p.dt = new
    p.dt_max = max(p.dt_max, new)


def _stop_tests(r, p, df, diag):
    """Get_f_ABC's tests after a block, in its order."""
    fit, A = diag.fitness, diag.amplitudes[0]
    if r.requiredFitness is not None and fit < r.requiredFitness:
        p.stop = Q.STOP_POOR_FIT
    elif A < r.minimumA:
        p.stop = Q.STOP_MIN_AMPLITUDE
    elif abs(df) <= r.fTol * abs(p.f) and abs(A - p.A) <= r.ATol * A:
        p.stop = Q.STOP_CONVERGED
    p.A = A


def find_q0star_batched(minimizer, temp, meanDensity, minPoints=None, record=None, precision=64):
    """Find_q0Star with each pass's q00 points evaluated as one batch."""
    def evaluate(qs):
        return evaluate_batched(minimizer, [(temp, meanDensity, q) for q in qs], precision=precision)
    return minimizer.Find_q0Star(temp, meanDensity, minPoints=minPoints, record=record, evaluate=evaluate)


def scan_density_batched(minimizer, temp, densities, checkpoint=None, precision=64, verbose=False):
    """scan_density with the q00 passes of all densities batched together:
    pass p of every density still scanning runs as one batch.  Results,
    statuses, q-scan rows and checkpoint files are scan_density's (wall_s is
    the time of the passes a density took part in); a checkpoint is written
    after every pass, and a resume skips the densities with a status."""
    qs_set = minimizer.qscan
    phi = np.asarray(list(densities), dtype=float)
    n = phi.size
    q0star, fmin, wall_s = [np.nan] * n, [np.nan] * n, [0.0] * n
    status, message = [''] * n, [''] * n
    rows, done = [], set()
    prev = Q._resume(checkpoint, phi) if checkpoint is not None else None
    if prev is not None:
        old, done, rows = prev
        for i in done:
            q0star[i], fmin[i], wall_s[i] = float(old['q0star'][i]), float(old['fmin'][i]), float(old['wall_s'][i])
            status[i], message[i] = str(old['status'][i]), str(old['message'][i])
    seed, noise = getattr(minimizer, 'seed', None), getattr(minimizer, 'noise', None)
    scanning = [i for i in range(n) if i not in done]
    q_best = {i: qs_set.q00_start for i in scanning}
    Deltaq = qs_set.Deltaq
    new_rows = {i: [] for i in scanning}
    for pass_idx in range(qs_set.passes):
        if not scanning:
            break
        t0 = time.perf_counter()
        grids = {i: Q.pass_grid(q_best[i], Deltaq, qs_set) for i in scanning}
        evals = _pass_evaluations(minimizer, temp, phi, scanning, grids, precision)
        wall = time.perf_counter() - t0
        still = []
        for i in scanning:
            wall_s[i] += wall
            if isinstance(evals[i], Exception):
                status[i], message[i] = Q.status_from_exception(evals[i])
                continue
            new_rows[i] += [Q.record_of(pass_idx, q, e) for q, e in zip(grids[i], evals[i])]
            try:
                q_best[i], spl = Q.pass_decision(grids[i], evals[i], minimizer.relax, qs_set.minPoints)
            except Exception as e:  # noqa: BLE001 - a density's failure is its status
                status[i], message[i] = Q.status_from_exception(e)
                continue
            if pass_idx == qs_set.passes - 1:
                q0star[i], fmin[i], status[i], message[i] = float(q_best[i]), float(spl(q_best[i])), Q.STATUS_OK, ''
            else:
                still.append(i)
        Deltaq *= qs_set.Deltaq_Refinement
        if verbose:
            print(f'pass {pass_idx}: {len(scanning)} densities in {wall:0.2f} s; {len(still)} continue')
        finished = [i for i in scanning if i not in still]
        scanning = still
        for i in finished:
            rows += [{'phi_index': i, 'pass_index': rec['pass'], **{k: v for k, v in rec.items() if k != 'pass'}}
                     for rec in new_rows.pop(i)]
        if checkpoint is not None and finished:
            Q._write_npz(os.path.join(checkpoint, Q.QSCAN_FILE), Q._qscan_table(rows, seed, noise))
            Q._write_npz(os.path.join(checkpoint, Q.FMIN_FILE),
                         Q._fmin_table(phi, q0star, fmin, status, message, wall_s))
    out = Q._fmin_table(phi, q0star, fmin, status, message, wall_s)
    out['qscan'] = Q._qscan_table(sorted(rows, key=lambda r: r['phi_index']), seed, noise)
    return out


def _pass_evaluations(minimizer, temp, phi, scanning, grids, precision):
    """{density index: its pass's Evaluations, or the exception that ended it}.
    All densities go in one batch; if that batch raises, each density is run
    as a batch of its own, so an error costs its density only (as in
    scan_density, where a density's exception becomes its status)."""
    points = [(temp, phi[i], q) for i in scanning for q in grids[i]]
    try:
        evals = evaluate_batched(minimizer, points, precision=precision)
    except Exception:  # noqa: BLE001 - retried density by density below
        out = {}
        for i in scanning:
            try:
                out[i] = evaluate_batched(minimizer, [(temp, phi[i], q) for q in grids[i]], precision=precision)
            except Exception as e:  # noqa: BLE001
                out[i] = e
        return out
    k = len(grids[scanning[0]]) if scanning else 0
    return {i: evals[j * k:(j + 1) * k] for j, i in enumerate(scanning)}
