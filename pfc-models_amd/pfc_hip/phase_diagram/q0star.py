"""Lattice scale q0* and fmin for one structure: the phase-diagram relaxation driver.

A restatement of the reference's model-agnostic driver
(phase_diagram/q0star.py: `MinimizeFreeEnergy._generateSim` :232-258,
`_takeStep` :261-307, `_adaptDt` :309-326, `Get_f_ABC` :329-406,
`Find_q0Star` :409-477), with its settings, stop codes, messages and
exceptions, over an adapter module with the reference's names (`build`,
`get_state`, `set_state`, `get_time`, `set_time`, `get_dt`, `step`,
`free_energy`, `diagnose`, `STRUCTURES`, `validate_*`).  An adapter may also
provide `run(sim, n)`, n steps in one call, which the driver then uses for a
block instead of n calls of `step`; the hip adapters do, so a block is one
asynchronous device call.

Per evaluation (`Get_f_ABC`): relax in blocks of `stepT`; a block whose free
energy rises by more than fTol |f| (or turns NaN) is rolled back and redone
as `divs` sub-blocks at dt/divs, recursively to `maxDepth`; stop when a block
changes f by <= fTol |f| and the primary amplitude A by <= ATol A
(converged), when the fitness falls below `requiredFitness`, when A falls
below `minimumA`, when the recursion is exhausted, or at totalT.  Optional
adaptive dt between sub-blocks (`dtMax`).  `Find_q0Star` scans q00 in passes
of `q00_Steps` points, splines f(q00) over the contiguous accepted run around
the argmin and takes the spline minimum nearest to it.
"""
from __future__ import annotations

import os
import time
from dataclasses import dataclass, field

import numpy as np
from scipy.interpolate import CubicSpline

MSG_INSUFFICIENT_POINTS = 'Insufficient points found that satisfy fitness constraints.'
MSG_NO_LOCAL_MINIMUM = 'No local minimum found for q00.'


class InsufficientPoints(Exception):
    """Find_q0Star: too few accepted q00 points."""


class NoLocalMinimum(Exception):
    """Find_q0Star: the spline of f(q00) has no usable minimum."""


STATUS_OK = 'ok'
STATUS_INSUFFICIENT_POINTS = 'insufficient_points'
STATUS_NO_LOCAL_MINIMUM = 'no_local_minimum'
STATUS_ERROR = 'error'

STOP_CONVERGED = 'converged'
STOP_MAX_DEPTH = 'max_depth'
STOP_NAN = 'nan'
STOP_POOR_FIT = 'poor_fit'
STOP_MIN_AMPLITUDE = 'min_amplitude'
STOP_TOTALT = 'totalT'
FAILED_STOPS = (STOP_MAX_DEPTH, STOP_NAN)


class MaxDepthReached(RuntimeError):
    """A block could not be completed at maxDepth; `nan`: the last failure was a NaN f."""

    def __init__(self, message, nan=False):
        super().__init__(message)
        self.nan = nan


@dataclass(frozen=True)
class RelaxSettings:
    """Relaxation by time stepping (the reference's RelaxSettings, q0star.py:124-149)."""
    dt: float
    totalT: float
    stepT: float
    divs: int
    maxDepth: int
    requiredFitness: float | None
    minimumA: float
    fTol: float = 1e-12
    ATol: float = 1e-6
    seed: int = 0
    noise: float = 1e-10
    dtMax: float | None = None
    dtRate: float = 1e-8
    dtGrow: float = 2.0
    dtChecks: int = 10


@dataclass(frozen=True)
class QScanSettings:
    """q00 scan with spline refinement (q0star.py:152-162)."""
    passes: int
    Deltaq: float
    q00_Steps: int
    Deltaq_Refinement: float
    q00_start: float = 1.0
    minPoints: int = 5


@dataclass(frozen=True)
class Diagnostics:
    """What an adapter's `diagnose` returns; the driver reads `fitness` and `amplitudes[0]`."""
    fitness: float
    amplitudes: tuple
    details: dict = field(default_factory=dict)


def status_from_exception(exc):
    """(status code, message) of an exception raised by Find_q0Star."""
    text = str(exc)
    if isinstance(exc, InsufficientPoints):
        return STATUS_INSUFFICIENT_POINTS, text
    if isinstance(exc, NoLocalMinimum):
        return STATUS_NO_LOCAL_MINIMUM, text
    return STATUS_ERROR, f'{type(exc).__name__}: {text}'


def _advance(ad, sim, n):
    run = getattr(ad, 'run', None)
    if run is not None:
        run(sim, n)
    else:
        for _ in range(n):
            ad.step(sim)


def noisy_state(state, seed, noise):
    """A fresh ansatz with each field multiplied by 1 + noise N(0, 1) from
    default_rng(seed) (one stream over the fields in order) and its mean
    restored (q0star.py:245-256)."""
    rng = np.random.default_rng(seed)
    fields = state if isinstance(state, tuple) else (state,)
    noisy = []
    for phi in fields:
        mean = phi.mean()
        phi = phi * (1.0 + noise * rng.standard_normal(phi.shape))
        noisy.append(phi + (mean - phi.mean()))
    return tuple(noisy) if isinstance(state, tuple) else noisy[0]


def add_ansatz_noise(ad, sim, seed, noise):
    ad.set_state(sim, noisy_state(ad.get_state(sim), seed, noise))


@dataclass
class Evaluation:
    """One Get_f_ABC: the values at the stop, why it stopped, its time steps,
    the largest dt it used and (optionally) the final state."""
    f: float
    fitness: float
    amplitudes: tuple
    stop: str
    steps: int
    dt_max: float
    state: object = None


def pass_decision(qs, rows, relax, minPoints):
    """One q00 pass of Find_q0Star (q0star.py:447-476): the accepted points
    (fitness >= requiredFitness unless the evaluation failed, A >= minimumA)
    must number at least minPoints overall and on the contiguous run around
    their argmin; the run is splined and the spline minimum nearest the argmin
    returned with the spline.  Raises InsufficientPoints / NoLocalMinimum."""
    f = np.array([r.f for r in rows], dtype=float)
    fit = np.array([np.nan if r.stop in FAILED_STOPS else r.fitness for r in rows], dtype=float)
    A = np.array([r.amplitudes[0] for r in rows], dtype=float)
    rf = -np.inf if relax.requiredFitness is None else relax.requiredFitness
    ok = (fit >= rf) & (A >= relax.minimumA)
    if qs[ok].size < minPoints:
        raise InsufficientPoints(MSG_INSUFFICIENT_POINTS)
    i_min = np.flatnonzero(ok)[np.argmin(f[ok])]
    q_min = qs[i_min]
    lo = hi = i_min
    while lo > 0 and ok[lo - 1]:
        lo -= 1
    while hi < ok.size - 1 and ok[hi + 1]:
        hi += 1
    if hi - lo + 1 < minPoints:
        raise InsufficientPoints(MSG_INSUFFICIENT_POINTS)
    spl = CubicSpline(qs[lo:hi + 1], f[lo:hi + 1])
    roots = spl.derivative().roots(extrapolate=False)
    roots = roots[spl.derivative(2)(roots) > 0]
    if roots.size == 0:
        raise NoLocalMinimum(MSG_NO_LOCAL_MINIMUM)
    return np.abs(min(roots, key=lambda x: abs(x - q_min))), spl


def pass_grid(q_best, Deltaq, qscan):
    """The q00 points of one pass: q00_Steps points spaced Deltaq about q_best."""
    half = qscan.q00_Steps // 2 * Deltaq
    return np.linspace(q_best - half, q_best + half, qscan.q00_Steps, endpoint=True)


class MinimizeFreeEnergy:
    def __init__(self, adapter, params, structure, cell, ansatz, relax, qscan, verbose=False):
        self.adapter = adapter
        self.params = adapter.validate_params(params)
        self.structure = adapter.STRUCTURES[structure] if isinstance(structure, str) else structure
        self.cell = adapter.validate_cell(cell)
        self.ansatz = adapter.validate_ansatz(ansatz)
        self.relax = relax
        self.qscan = qscan
        # every setting is also an attribute of the minimiser (dt, totalT, ...,
        # passes, Deltaq, ...), the names callers of the reference read
        for settings in (relax, qscan):
            for name in settings.__dataclass_fields__:
                setattr(self, name, getattr(settings, name))
        self.last_stop, self.last_steps, self.last_dt_max = None, 0, float('nan')
        self.last_state = None
        if verbose:
            print(f'model {adapter.SLUG} {self.params}; structure {getattr(self.structure, "name", self.structure)}, '
                  f'cell {self.cell}; ansatz {self.ansatz}; {relax}; {qscan}')

    # -- one simulation ------------------------------------------------------
    def _generateSim(self, temp, meanDensity, q00, dt, densityField=None, verbose=False):
        sim = self.adapter.build(self.params, self.structure, self.cell, self.ansatz, temp, meanDensity, q00, dt,
                                 state=densityField, nThreads=1)
        if densityField is None and self.relax.noise:
            add_ansatz_noise(self.adapter, sim, self.relax.seed, self.relax.noise)
        return sim

    def _takeStep(self, sim, temp, meanDensity, q00, stepT, divs=10, depth=0, maxDepth=None, verbose=False):
        """One block of stepT; rolled back and refined on a rise of f (or NaN).
        Returns the block's change of f."""
        if maxDepth is None:
            maxDepth = self.relax.maxDepth
        ad = self.adapter
        dt = ad.get_dt(sim)
        n = int(round(stepT / dt))
        before, t_before, f_before = ad.get_state(sim), ad.get_time(sim), ad.free_energy(sim)
        _advance(ad, sim, n)
        self.last_steps += n
        f = ad.free_energy(sim)
        df = f - f_before
        if not (np.isnan(f) or df > self.relax.fTol * abs(f_before)):
            return df
        return self._rollback(sim, temp, meanDensity, q00, stepT, before, t_before, f_before, bool(np.isnan(f)),
                              divs=divs, depth=depth, maxDepth=maxDepth, verbose=verbose)

    def _rollback(self, sim, temp, meanDensity, q00, stepT, before, t_before, f_before, nan, divs=10, depth=0,
                  maxDepth=None, verbose=False):
        """A block that failed: restore `before` at t_before and redo it as
        `divs` sub-blocks at dt / divs, or raise MaxDepthReached at maxDepth.
        Returns the block's change of f."""
        if maxDepth is None:
            maxDepth = self.relax.maxDepth
        ad = self.adapter
        dt = ad.get_dt(sim)
        ad.set_state(sim, before)
        ad.set_time(sim, t_before)
        if depth >= maxDepth:
            raise MaxDepthReached(f'Max recurrsion depth reached at t={ad.get_time(sim):0.6f}', nan=nan)
        sub = self._generateSim(temp, meanDensity, q00, dt / divs, ad.get_state(sim), verbose=verbose)
        ad.set_time(sub, ad.get_time(sim))
        for _ in range(divs):
            self._takeStep(sub, temp, meanDensity, q00, stepT / divs, divs=divs, depth=depth + 1, maxDepth=maxDepth,
                           verbose=verbose)
            ad.set_state(sim, ad.get_state(sub))
            ad.set_time(sim, ad.get_time(sub))
        return ad.free_energy(sim) - f_before

    def _adaptDt(self, sim, temp, meanDensity, q00, stepT, delta_f, f, A, A_last, verbose=False):
        """Next block's dt from the block's relative rate; the simulation is rebuilt only on a change."""
        ad = self.adapter
        r = self.relax
        dt = ad.get_dt(sim)
        rate = max(abs(delta_f) / max(abs(f), 1e-300), abs(A - A_last) / max(A, 1e-300)) / stepT
        target = r.dtMax / np.sqrt(1.0 + (rate / r.dtRate) ** 2)
        target = min(max(target, r.dt), r.dtMax, dt * r.dtGrow)
        n = int(np.ceil(stepT / target - 1e-9))
        new = stepT / n
        if abs(new - dt) <= 1e-12 * dt:
            return sim
        t = ad.get_time(sim)
        sim = self._generateSim(temp, meanDensity, q00, new, ad.get_state(sim), verbose=verbose)
        ad.set_time(sim, t)
        self.last_dt_max = max(self.last_dt_max, new)
        return sim

    # -- one evaluation --------------------------------------------------------
    def Get_f_ABC(self, temp, meanDensity, q00, maxDepth=None, verbose=False, state=None):
        """Relax at fixed box scale q00; returns (f, fitness, *amplitudes) of
        the state at the stop, which is left in `last_stop` (and the step
        count in `last_steps`, the final field in `last_state`)."""
        if maxDepth is None:
            maxDepth = self.relax.maxDepth
        ad, r = self.adapter, self.relax
        blocks = int(r.totalT / r.stepT)
        dt0 = r.stepT / int(np.ceil(r.stepT / r.dt))  # keeps stepT exact
        sim = self._generateSim(temp, meanDensity, q00, dt0, densityField=state, verbose=verbose)
        f_last = ad.free_energy(sim)
        A_last = ad.diagnose(sim, self.structure, self.cell).amplitudes[0]
        self.last_stop, self.last_steps = STOP_TOTALT, 0
        self.last_dt_max = ad.get_dt(sim)
        for _ in range(blocks):
            try:
                if r.dtMax is None:
                    df = self._takeStep(sim, temp, meanDensity, q00, r.stepT, divs=r.divs, maxDepth=maxDepth,
                                        verbose=verbose)
                else:
                    f0, sub_T = f_last, r.stepT / r.dtChecks
                    A_s = A_last
                    for k in range(r.dtChecks):
                        d_sub = self._takeStep(sim, temp, meanDensity, q00, sub_T, divs=r.divs, maxDepth=maxDepth,
                                               verbose=verbose)
                        f_new = ad.free_energy(sim)
                        A_new = ad.diagnose(sim, self.structure, self.cell).amplitudes[0]
                        if k < r.dtChecks - 1:
                            sim = self._adaptDt(sim, temp, meanDensity, q00, sub_T, d_sub, f_new, A_new, A_s, verbose)
                        A_s = A_new
                    df = ad.free_energy(sim) - f0
            except MaxDepthReached as e:
                self.last_stop = STOP_NAN if e.nan else STOP_MAX_DEPTH
                break
            f_last = ad.free_energy(sim)
            diag = ad.diagnose(sim, self.structure, self.cell)
            fit, A = diag.fitness, diag.amplitudes[0]
            if r.requiredFitness is not None and fit < r.requiredFitness:
                self.last_stop = STOP_POOR_FIT
                break
            if A < r.minimumA:
                self.last_stop = STOP_MIN_AMPLITUDE
                break
            if abs(df) <= r.fTol * abs(f_last) and abs(A - A_last) <= r.ATol * A:
                self.last_stop = STOP_CONVERGED
                break
            if r.dtMax is not None:
                sim = self._adaptDt(sim, temp, meanDensity, q00, r.stepT / r.dtChecks, df / r.dtChecks, f_last, A,
                                    A_last, verbose)
            A_last = A
        f = ad.free_energy(sim)
        diag = ad.diagnose(sim, self.structure, self.cell)
        self.last_state = ad.get_state(sim)
        return (f, diag.fitness, *diag.amplitudes)

    def evaluate(self, temp, meanDensity, q00, verbose=False, state=None):
        """Get_f_ABC packed as an Evaluation (final state included)."""
        out = self.Get_f_ABC(temp=temp, meanDensity=meanDensity, q00=q00, verbose=verbose, state=state)
        return Evaluation(float(out[0]), float(out[1]), tuple(float(a) for a in out[2:]), self.last_stop,
                          self.last_steps, self.last_dt_max, self.last_state)

    # -- the q00 scan ------------------------------------------------------------
    def Find_q0Star(self, temp, meanDensity, minPoints=None, verbose=False, record=None, evaluate=None):
        """q00 minimising f at (temp, meanDensity); returns (q0*, fmin).
        `evaluate(q00s)`, if given, evaluates a whole pass at once and returns
        one Evaluation per point (the batched form); by default each point is a
        Get_f_ABC call.  `record`, if given, receives one dict per evaluation
        (pass, q00, f, fitness, amplitudes, stop, steps, dt_max)."""
        q = self.qscan
        if minPoints is None:
            minPoints = q.minPoints
        Deltaq = q.Deltaq
        q_best = q.q00_start
        spl = None
        for pass_idx in range(q.passes):
            qs = pass_grid(q_best, Deltaq, q)
            if evaluate is not None:
                rows = list(evaluate(qs))
                for q00, r in zip(qs, rows):
                    if record is not None:
                        record(record_of(pass_idx, q00, r))
            else:
                rows = []
                for q00 in qs:  # recorded as they finish: an exception keeps the earlier records
                    rows.append(self.evaluate(temp, meanDensity, q00, verbose=verbose))
                    if record is not None:
                        record(record_of(pass_idx, q00, rows[-1]))
            q_best, spl = pass_decision(qs, rows, self.relax, minPoints)
            Deltaq *= q.Deltaq_Refinement
        return q_best, spl(q_best)


def record_of(pass_idx, q00, r):
    """The q-scan record of one evaluation."""
    return {'pass': pass_idx, 'q00': float(q00), 'f': float(r.f), 'fitness': float(r.fitness),
            'amplitudes': tuple(float(a) for a in r.amplitudes), 'stop': r.stop, 'steps': r.steps,
            'dt_max': r.dt_max}


# ---------------------------------------------------------------------------
# a density list: one structure at one temperature (q0star.py:491-656)
# ---------------------------------------------------------------------------
FMIN_FILE = 'fmin.npz'
QSCAN_FILE = 'qscan.npz'
_ROW_DEFAULTS = {'stop': '', 'steps': -1, 'dt_max': float('nan')}


def _write_npz(path, arrays):
    """Write beside the target, then rename over it: a reader never sees half a file."""
    part = path + '.tmp'
    with open(part, 'wb') as fh:
        np.savez(fh, **arrays)
    os.replace(part, path)


def _fmin_table(phi, q0star, fmin, status, message, wall_s):
    return {'phi': np.asarray(phi, dtype=float), 'q0star': np.asarray(q0star, dtype=float),
            'fmin': np.asarray(fmin, dtype=float), 'status': np.asarray(status, dtype=str),
            'message': np.asarray(message, dtype=str), 'wall_s': np.asarray(wall_s, dtype=float)}


def _qscan_table(rows, seed, noise):
    """One row per Get_f_ABC evaluation; amplitudes NaN-padded to the widest row."""
    width = max((len(r['amplitudes']) for r in rows), default=0)
    amps = np.full((len(rows), width), np.nan)
    for j, r in enumerate(rows):
        amps[j, :len(r['amplitudes'])] = r['amplitudes']

    def col(key, dtype):
        return np.asarray([r.get(key, _ROW_DEFAULTS.get(key)) for r in rows], dtype=dtype)
    return {'phi_index': col('phi_index', np.int64), 'pass_index': col('pass_index', np.int64),
            'q00': col('q00', float), 'f': col('f', float), 'fitness': col('fitness', float), 'amplitudes': amps,
            'stop': col('stop', str), 'steps': col('steps', np.int64), 'dt_max': col('dt_max', float),
            'seed': np.asarray(-1 if seed is None else seed, dtype=np.int64),
            'noise': np.asarray(np.nan if noise is None else noise, dtype=float)}


def _resume(checkpoint, phi):
    """(previous fmin arrays, indices with a status, their q-scan rows), or None
    without a checkpoint.  A different density list is refused."""
    path = os.path.join(checkpoint, FMIN_FILE)
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        old = {k: z[k] for k in z.files}
    if old['phi'].shape != phi.shape or not np.array_equal(old['phi'], phi):
        raise ValueError(f'{path} holds a different density list; refusing to resume into it')
    done = {i for i, s in enumerate(old['status']) if str(s)}
    rows = []
    path = os.path.join(checkpoint, QSCAN_FILE)
    if os.path.exists(path):
        with np.load(path, allow_pickle=False) as z:
            q = {k: z[k] for k in z.files}
        for j in np.flatnonzero(np.isin(q['phi_index'], sorted(done))):
            a = q['amplitudes'][j]
            row = {'phi_index': int(q['phi_index'][j]), 'pass_index': int(q['pass_index'][j]),
                   'q00': float(q['q00'][j]), 'f': float(q['f'][j]), 'fitness': float(q['fitness'][j]),
                   'amplitudes': tuple(float(x) for x in a[~np.isnan(a)])}
            for key, default in _ROW_DEFAULTS.items():
                row[key] = type(default)(q[key][j]) if key in q else default
            rows.append(row)
    return old, done, rows


def scan_density(minimizer, temp, densities, checkpoint=None, verbose=False):
    """Find_q0Star at each density of a list.  A failure costs its density
    only: status 'ok', 'insufficient_points', 'no_local_minimum' (message the
    exception's) or 'error' ('Type: text'); q0star and fmin are NaN unless ok.
    Every evaluation is kept as a q-scan row (phi_index, pass_index, q00, f,
    fitness, amplitudes, stop, steps, dt_max) with the ansatz-noise seed and
    amplitude.  With `checkpoint` (an existing directory) qscan.npz, then
    fmin.npz, are rewritten after each density, and a restart skips the
    densities that already have a status.  Returns the fmin arrays plus
    'qscan', the q-scan arrays."""
    phi = np.asarray(list(densities), dtype=float)
    n = phi.size
    q0star, fmin, wall_s = [np.nan] * n, [np.nan] * n, [np.nan] * n
    status, message = [''] * n, [''] * n
    rows, done = [], set()
    prev = _resume(checkpoint, phi) if checkpoint is not None else None
    if prev is not None:
        old, done, rows = prev
        for i in done:
            q0star[i], fmin[i], wall_s[i] = float(old['q0star'][i]), float(old['fmin'][i]), float(old['wall_s'][i])
            status[i], message[i] = str(old['status'][i]), str(old['message'][i])
    seed, noise = getattr(minimizer, 'seed', None), getattr(minimizer, 'noise', None)
    for i in range(n):
        if i in done:
            continue
        recs = []
        t0 = time.perf_counter()
        try:
            q, f = minimizer.Find_q0Star(temp=temp, meanDensity=phi[i], verbose=verbose, record=recs.append)
            q0star[i], fmin[i], status[i], message[i] = float(q), float(f), STATUS_OK, ''
        except Exception as e:  # noqa: BLE001 - a density's failure is its status
            status[i], message[i] = status_from_exception(e)
        wall_s[i] = time.perf_counter() - t0
        rows += [{'phi_index': i, 'pass_index': r['pass'], 'q00': r['q00'], 'f': r['f'], 'fitness': r['fitness'],
                  'amplitudes': r['amplitudes'], **{k: r.get(k, d) for k, d in _ROW_DEFAULTS.items()}} for r in recs]
        if verbose:
            print(f'phi={phi[i]}: {status[i]} q0*={q0star[i]} fmin={fmin[i]} ({wall_s[i]:0.1f} s) {message[i]}')
        if checkpoint is not None:
            # rows first: rows of a density without a status are dropped on resume
            _write_npz(os.path.join(checkpoint, QSCAN_FILE), _qscan_table(rows, seed, noise))
            _write_npz(os.path.join(checkpoint, FMIN_FILE), _fmin_table(phi, q0star, fmin, status, message, wall_s))
    out = _fmin_table(phi, q0star, fmin, status, message, wall_s)
    out['qscan'] = _qscan_table(rows, seed, noise)
    return out
