"""The phase-diagram relaxation driver (pfc_hip.phase_diagram.q0star) on a
toy adapter: stop reasons, rollback refinement, the amplitude guard, adaptive
dt, the q00 scan's decisions and scan_density's statuses and checkpoints.
The behaviours are those the reference's driver tests pin
(tests/test_phase_diagram_q0star.py:214-470, 674-699); the toy, its regimes
and the crafted f(q00) tables here are this suite's own.

The toy's state is a pair of arrays (the driver must treat it as opaque)
decaying by explicit Euler steps towards zero; f = floor + mean(x^2) +
mean(y^2), so a relaxed evaluation returns the floor, which the regime sets
as a function of (temp, meanDensity, q00).
"""
import dataclasses
import inspect

import numpy as np
import pytest

from pfc_hip.phase_diagram import q0star as Q

QSTAR = 0.97          # the toy's lattice scale
CURV = 2.0            # f_floor = temp + rho^2 + CURV (q00 - QSTAR)^2
T_BREAK = 4.0         # unstable regimes break once t >= T_BREAK
DT_OK = 0.005         # the 'stiff' regime is stable at dt <= DT_OK


class _Sim:
    def __init__(self, floor, fit, amp, dt, x, y, kind):
        self.floor, self.fit, self.amp, self.dt = floor, fit, amp, dt
        self.x, self.y, self.t, self.kind = x, y, 0.0, kind


class Toy:
    """Adapter with only the names the driver calls (plus `run` when
    `batched_run`).  `KINDS` maps a density to a regime:
      relax    f_floor = temp + rho^2 + CURV (q00 - QSTAR)^2, fitness 1 - (q00 - QSTAR)^2
      slope    f_floor = temp + q00 (no minimum in q00)
      misfit   fitness 0.4 at every q00
      stiff    NaN once t >= T_BREAK unless dt <= DT_OK
      blowup   NaN once t >= T_BREAK at any dt
      grow     state 0, amplitude min(0.25, 2e-6 e^t)
      raise    step raises RuntimeError once t >= T_BREAK
      reject   build raises ValueError
    `table=(f, fitness, A)` (callables of q00) replaces the regimes with a
    stationary state: each evaluation converges after one block."""
    SLUG = 'toy'
    STRUCTURES = {'dots': 'dots'}
    KINDS = {0.5: 'relax', 1.5: 'slope', 2.5: 'misfit', 0.7: 'stiff', 0.8: 'blowup', 0.6: 'grow', 0.3: 'raise',
             -1.0: 'reject'}

    def __init__(self, table=None, batched_run=False, stop_at=None):
        self.table, self.stop_at = table, stop_at
        self.builds = self.from_state = self.steps = 0
        self.dts = []
        if batched_run:
            self.run = self._run

    def validate_params(self, p):
        return dict(p)

    validate_cell = validate_ansatz = validate_params

    def build(self, params, structure, cell, ansatz, temp, meanDensity, q00, dt, state=None, nThreads=1):
        kind = self.KINDS.get(float(meanDensity), 'relax')
        if self.stop_at is not None and float(meanDensity) == self.stop_at:
            raise KeyboardInterrupt
        if kind == 'reject':
            raise ValueError('density out of range')
        self.builds += 1
        self.dts.append(dt)
        amp = 0.25
        if self.table is not None:
            floor, fit, amp = self.table[0](q00), self.table[1](q00), self.table[2](q00)
        elif kind == 'slope':
            floor, fit = temp + q00, 1.0
        elif kind == 'misfit':
            floor, fit = temp + (q00 - QSTAR) ** 2, 0.4
        else:
            floor = temp + meanDensity ** 2 + CURV * (q00 - QSTAR) ** 2
            fit = 1.0 - (q00 - QSTAR) ** 2
        if state is not None:
            assert isinstance(state, tuple) and len(state) == 2
            self.from_state += 1
            x, y = state[0].copy(), state[1].copy()
        elif self.table is not None or kind == 'grow':
            x, y = np.zeros(4), np.zeros(6)
        else:
            x, y = np.linspace(0.2, 0.6, 4), np.full(6, -0.3)
        return _Sim(floor, fit, amp, dt, x, y, kind)

    def get_state(self, s):
        return (s.x.copy(), s.y.copy())

    def set_state(self, s, state):
        s.x[...], s.y[...] = state

    def get_time(self, s):
        return s.t

    def set_time(self, s, t):
        s.t = t

    def get_dt(self, s):
        return s.dt

    def step(self, s):
        self.steps += 1
        s.x *= 1 - 0.4 * s.dt
        s.y *= 1 - 0.9 * s.dt
        s.t += s.dt
        if s.t >= T_BREAK:
            if s.kind == 'blowup' or (s.kind == 'stiff' and s.dt > DT_OK):
                s.x[...] = np.nan
            if s.kind == 'raise':
                raise RuntimeError('toy integrator failed')

    def _run(self, s, n):
        for _ in range(n):
            self.step(s)

    def free_energy(self, s):
        return s.floor + np.mean(s.x ** 2) + np.mean(s.y ** 2)

    def diagnose(self, s, structure, cell):
        a = min(0.25, 2e-6 * np.exp(s.t)) if s.kind == 'grow' else s.amp
        return Q.Diagnostics(fitness=s.fit, amplitudes=(a, 0.5 * a))


RELAX = Q.RelaxSettings(dt=0.1, totalT=2e3, stepT=10.0, divs=10, maxDepth=3, requiredFitness=0.995, minimumA=1e-6)
QSCAN = Q.QScanSettings(passes=3, Deltaq=0.02, q00_Steps=21, Deltaq_Refinement=0.05)


def _mini(ad, relax=RELAX, qscan=QSCAN):
    return Q.MinimizeFreeEnergy(ad, {}, 'dots', {}, {}, relax, qscan)


def _floor(temp, rho, q00):
    return temp + rho ** 2 + CURV * (q00 - QSTAR) ** 2


def test_driver_names_no_model():
    src = inspect.getsource(Q)
    for token in ('.State', 'sim.n', 'sim.phi', 'kvector', 'pfc_hip.models', 'from ..models', 'Geometry'):
        assert token not in src, token


def test_settings_defaults_and_attributes():
    r = Q.RelaxSettings(dt=0.1, totalT=1e6, stepT=1e2, divs=10, maxDepth=4, requiredFitness=0.99, minimumA=1e-6)
    assert (r.fTol, r.ATol, r.seed, r.noise, r.dtMax) == (1e-12, 1e-6, 0, 1e-10, None)
    assert Q.QScanSettings(passes=1, Deltaq=0.02, q00_Steps=21, Deltaq_Refinement=0.05).minPoints == 5
    m = _mini(Toy())
    assert (m.dt, m.stepT, m.maxDepth, m.noise, m.passes, m.q00_Steps, m.minPoints) == (0.1, 10.0, 3, 1e-10, 3, 21, 5)


def test_converges_to_the_floor_in_whole_blocks():
    ad = Toy()
    m = _mini(ad)
    f, fit, A, B = m.Get_f_ABC(temp=0.2, meanDensity=0.5, q00=1.0)
    assert m.last_stop == Q.STOP_CONVERGED
    assert f == pytest.approx(_floor(0.2, 0.5, 1.0), abs=1e-11)
    assert (fit, A, B) == (1.0 - (1.0 - QSTAR) ** 2, 0.25, 0.125)
    assert m.last_steps == ad.steps and m.last_steps % 100 == 0 and m.last_steps > 100
    assert ad.builds == 1 and ad.from_state == 0 and m.last_state[0].shape == (4,)


def test_run_entry_point_gives_the_same_evaluation():
    a, b = Toy(), Toy(batched_run=True)
    ra = _mini(a).Get_f_ABC(temp=0.2, meanDensity=0.7, q00=1.0)
    rb = _mini(b).Get_f_ABC(temp=0.2, meanDensity=0.7, q00=1.0)
    assert ra == rb and a.steps == b.steps and a.dts == b.dts


def test_rollback_refines_dt_from_the_opaque_state():
    # the stiff regime goes NaN at dt = 0.1 and 0.01, and is stable at 0.001:
    # the rollback must rebuild at dt/divs twice, from the (x, y) tuple
    ad = Toy()
    m = _mini(ad)
    f, *_ = m.Get_f_ABC(temp=0.2, meanDensity=0.7, q00=1.0)
    assert m.last_stop == Q.STOP_CONVERGED
    assert f == pytest.approx(_floor(0.2, 0.7, 1.0), abs=1e-11)
    assert ad.from_state >= 2 and min(ad.dts) == pytest.approx(1e-3)


def test_nan_at_every_dt_stops_with_nan_and_is_rejected():
    ad = Toy()
    m = _mini(ad)
    f, *_ = m.Get_f_ABC(temp=0.2, meanDensity=0.8, q00=1.0)
    assert m.last_stop == Q.STOP_NAN and np.isfinite(f)
    assert min(ad.dts) == pytest.approx(RELAX.dt / RELAX.divs ** RELAX.maxDepth)
    r = Q.scan_density(_mini(Toy()), 0.2, [0.8])
    assert str(r['status'][0]) == Q.STATUS_INSUFFICIENT_POINTS
    assert set(r['qscan']['stop']) == {Q.STOP_NAN}


def test_rise_without_nan_at_max_depth_is_max_depth():
    # a block whose f rises at every dt (no NaN): floor raised by each step
    class Rising(Toy):
        def step(self, s):
            super().step(s)
            s.floor += 1e-6
    ad = Rising()
    m = _mini(ad)
    m.Get_f_ABC(temp=0.2, meanDensity=0.5, q00=1.0)
    assert m.last_stop == Q.STOP_MAX_DEPTH


def test_amplitude_guard_waits_for_the_growing_seed():
    # f is constant; the amplitude grows from 2e-6 and saturates at t = 11.7:
    # converged after the block in which A stops changing (t = 30)
    m = _mini(Toy())
    f, fit, A, B = m.Get_f_ABC(temp=0.2, meanDensity=0.6, q00=1.0)
    assert m.last_stop == Q.STOP_CONVERGED and A == 0.25 and m.last_steps == 300


def test_poor_fit_and_min_amplitude_stops():
    m = _mini(Toy())
    m.Get_f_ABC(temp=0.2, meanDensity=2.5, q00=1.0)
    assert m.last_stop == Q.STOP_POOR_FIT and m.last_steps == 100
    m = _mini(Toy(table=(lambda q: 0.0, lambda q: 1.0, lambda q: 5e-7)))
    m.Get_f_ABC(temp=0.0, meanDensity=0.5, q00=1.0)
    assert m.last_stop == Q.STOP_MIN_AMPLITUDE


def test_total_time_stop():
    relax = dataclasses.replace(RELAX, totalT=20.0)
    m = _mini(Toy(), relax)
    m.Get_f_ABC(temp=0.2, meanDensity=0.5, q00=1.0)
    assert m.last_stop == Q.STOP_TOTALT and m.last_steps == 200


def test_other_errors_propagate():
    with pytest.raises(RuntimeError, match='toy integrator failed'):
        _mini(Toy()).Get_f_ABC(temp=0.2, meanDensity=0.3, q00=1.0)


def test_required_fitness_none():
    relax = dataclasses.replace(RELAX, requiredFitness=None)
    m = _mini(Toy(), relax)
    m.Get_f_ABC(temp=0.2, meanDensity=2.5, q00=1.0)       # fitness 0.4 is no stop now
    assert m.last_stop == Q.STOP_CONVERGED
    q, _ = m.Find_q0Star(temp=0.2, meanDensity=0.5)
    assert q == pytest.approx(QSTAR, abs=1e-9)


def test_ansatz_noise_is_seeded_and_keeps_the_mean():
    class Noisy(Toy):
        def build(self, *a, **k):
            s = super().build(*a, **k)
            if k.get('state') is None:
                s.x, s.y = np.linspace(0.2, 0.6, 64), np.full(36, -0.3)
            return s
    states = []
    for seed in (0, 0, 1):
        m = _mini(Noisy(), dataclasses.replace(RELAX, seed=seed, noise=1e-3))
        s = m._generateSim(0.2, 0.5, 1.0, 0.1)
        states.append((s.x.copy(), s.y.copy()))
    assert all((a == b).all() for a, b in zip(states[0], states[1]))
    assert not (states[0][0] == states[2][0]).all()
    for x, y in states:
        assert x.mean() == pytest.approx(0.4, abs=1e-15) and y.mean() == pytest.approx(-0.3, abs=1e-15)
        assert 0 < np.abs(x / np.linspace(0.2, 0.6, 64) - 1).max() < 1e-2
    m = _mini(Noisy(), dataclasses.replace(RELAX, noise=0.0))
    s = m._generateSim(0.2, 0.5, 1.0, 0.1)
    assert (s.x == np.linspace(0.2, 0.6, 64)).all()


def test_find_q0star_known_minimum_and_records():
    ad = Toy()
    m = _mini(ad)
    recs = []
    q, fmin = m.Find_q0Star(temp=0.2, meanDensity=0.5, record=recs.append)
    assert q == pytest.approx(QSTAR, abs=1e-9)
    assert float(fmin) == pytest.approx(0.2 + 0.25, abs=1e-11)
    assert len(recs) == QSCAN.passes * QSCAN.q00_Steps
    assert sorted(recs[0]) == ['amplitudes', 'dt_max', 'f', 'fitness', 'pass', 'q00', 'steps', 'stop']
    assert [r['pass'] for r in recs[::QSCAN.q00_Steps]] == [0, 1, 2]
    assert recs[0]['amplitudes'] == (0.25, 0.125)
    # pass 0 spans 0.80..1.20: fitness 1 - (q00 - 0.97)^2 < 0.995 beyond 0.07 from q*
    assert all(r['stop'] == (Q.STOP_POOR_FIT if r['fitness'] < 0.995 else Q.STOP_CONVERGED) for r in recs)
    assert sum(r['stop'] == Q.STOP_POOR_FIT for r in recs) > 0
    assert sum(r['steps'] for r in recs) == ad.steps
    q2, f2 = _mini(Toy()).Find_q0Star(temp=0.2, meanDensity=0.5)
    assert (q2, float(f2)) == (q, float(fmin))


def test_status_from_exception():
    assert Q.status_from_exception(Q.InsufficientPoints(Q.MSG_INSUFFICIENT_POINTS)) == \
        (Q.STATUS_INSUFFICIENT_POINTS, Q.MSG_INSUFFICIENT_POINTS)
    assert Q.status_from_exception(Q.NoLocalMinimum(Q.MSG_NO_LOCAL_MINIMUM)) == \
        (Q.STATUS_NO_LOCAL_MINIMUM, Q.MSG_NO_LOCAL_MINIMUM)
    # by type only: the message of a plain exception does not make it a status
    assert Q.status_from_exception(ValueError(Q.MSG_NO_LOCAL_MINIMUM)) == \
        (Q.STATUS_ERROR, 'ValueError: ' + Q.MSG_NO_LOCAL_MINIMUM)
    assert Q.status_from_exception(Q.MaxDepthReached('x'))[0] == Q.STATUS_ERROR


def test_scan_density_statuses_and_rows():
    r = Q.scan_density(_mini(Toy()), 0.2, [0.5, 1.5, 2.5, -1.0, 0.3])
    assert list(r['status']) == ['ok', 'no_local_minimum', 'insufficient_points', 'error', 'error']
    assert list(r['message']) == ['', Q.MSG_NO_LOCAL_MINIMUM, Q.MSG_INSUFFICIENT_POINTS,
                                  'ValueError: density out of range', 'RuntimeError: toy integrator failed']
    assert r['q0star'][0] == pytest.approx(QSTAR, abs=1e-9)
    assert np.isnan(r['q0star'][1:]).all() and np.isnan(r['fmin'][1:]).all() and np.isfinite(r['wall_s']).all()
    qs = r['qscan']
    assert np.bincount(qs['phi_index'], minlength=5).tolist() == [63, 21, 21, 0, 0]
    assert qs['amplitudes'].shape == (105, 2)
    assert set(qs) == {'phi_index', 'pass_index', 'q00', 'f', 'fitness', 'amplitudes', 'stop', 'steps', 'dt_max',
                       'seed', 'noise'}
    assert (int(qs['seed']), float(qs['noise'])) == (0, 1e-10)
    assert set(qs['stop']) == {Q.STOP_CONVERGED, Q.STOP_POOR_FIT} and (qs['steps'] > 0).all()


def test_scan_density_checkpoint_and_resume(tmp_path):
    rhos = [0.5, 1.5, 0.45, 2.5]
    fresh = Q.scan_density(_mini(Toy()), 0.2, rhos)
    with pytest.raises(KeyboardInterrupt):
        Q.scan_density(_mini(Toy(stop_at=0.45)), 0.2, rhos, checkpoint=str(tmp_path))
    with np.load(tmp_path / Q.FMIN_FILE) as z:
        assert list(z['status']) == ['ok', 'no_local_minimum', '', '']
    assert sorted(p.name for p in tmp_path.iterdir()) == [Q.FMIN_FILE, Q.QSCAN_FILE]
    ad = Toy()
    resumed = Q.scan_density(_mini(ad), 0.2, rhos, checkpoint=str(tmp_path))
    assert ad.builds - ad.from_state == 4 * QSCAN.q00_Steps          # 3 passes + 1 pass
    for k in ('phi', 'q0star', 'fmin', 'status', 'message'):
        np.testing.assert_array_equal(resumed[k], fresh[k])
    with np.load(tmp_path / Q.QSCAN_FILE) as z:
        assert z['q00'].size == fresh['qscan']['q00'].size
        np.testing.assert_array_equal(np.sort(z['f']), np.sort(fresh['qscan']['f']))
        np.testing.assert_array_equal(np.sort(z['stop']), np.sort(fresh['qscan']['stop']))
    ad = Toy()
    Q.scan_density(_mini(ad), 0.2, rhos, checkpoint=str(tmp_path))
    assert ad.builds == 0
    with pytest.raises(ValueError):
        Q.scan_density(_mini(Toy()), 0.2, [0.5], checkpoint=str(tmp_path))


# -- the q00 scan's decisions on crafted f(q00) tables (pass 0 spans 0.80..1.20 by 0.02)
def _table_find(f, fit, A=lambda q: 0.25, record=None):
    return _mini(Toy(table=(f, fit, A))).Find_q0Star(temp=0.0, meanDensity=0.5, record=record)


def _table_status(f, fit, A=lambda q: 0.25):
    r = Q.scan_density(_mini(Toy(table=(f, fit, A))), 0.0, [0.5])
    return str(r['status'][0])


def test_argmin_is_read_in_the_accepted_range():
    # deepest minimum at 1.10 (unfiltered index 15), a shallower one at 0.94;
    # 0.80..0.88 rejected (5 points).  Indexing the full range with the
    # filtered argmin would start from 1.00, nearer the shallow minimum.
    def f(q):
        return -np.exp(-((q - 1.10) / 0.035) ** 2) - 0.6 * np.exp(-((q - 0.94) / 0.035) ** 2)
    q, _ = _table_find(f, lambda q: 0.3 if q < 0.89 else 1.0)
    assert q == pytest.approx(1.10, abs=1e-4)


def test_only_spline_minima_are_candidates():
    # f = -(q - 0.86)^2 (q - 1.16): minimum at 0.86, maximum at 1.06, grid argmin
    # at the edge 1.20, whose nearest stationary point is the maximum
    q, _ = _table_find(lambda q: -(q - 0.86) ** 2 * (q - 1.16), lambda q: 1.0)
    assert q == pytest.approx(0.86, abs=1e-9)


def test_one_acceptance_mask_for_count_argmin_and_spline():
    # crystal on 0.87..1.13 with f = 2 (q - 0.99)^2; outside, amplitude 5e-7 <
    # minimumA at f = -5e-4, below every crystal point
    def crystal(q):
        return 0.87 < q < 1.13
    q, _ = _table_find(lambda q: 2 * (q - 0.99) ** 2 if crystal(q) else -5e-4, lambda q: 1.0,
                       lambda q: 0.25 if crystal(q) else 5e-7)
    assert q == pytest.approx(0.99, abs=1e-9)


def _two_runs(q):
    return 0.5 * (q - 0.96) ** 2 - 0.02 if q < 0.91 else 0.5 * (q - 1.08) ** 2


def test_spline_stays_on_the_run_around_the_argmin():
    # run 0.80..0.90 still falling at its edge (the argmin), a rejected gap
    # 0.92..0.98, a run 1.00..1.20 with a minimum at 1.08 above the argmin:
    # no minimum on the argmin's run, decided in pass 0
    recs = []
    with pytest.raises(Q.NoLocalMinimum):
        _table_find(_two_runs, lambda q: 0.3 if 0.91 < q < 0.99 else 1.0, record=recs.append)
    assert {r['pass'] for r in recs} == {0}
    # minPoints applies to that run: 3 points (0.86..0.90) with 14 accepted overall
    assert _table_status(_two_runs, lambda q: 0.3 if (q < 0.85 or 0.91 < q < 0.99) else 1.0) == \
        Q.STATUS_INSUFFICIENT_POINTS


def test_min_points_means_at_least():
    # exactly minPoints = 5 points on the argmin's run (0.82..0.90) is enough to
    # proceed (and then find no minimum on it)
    assert _table_status(_two_runs, lambda q: 0.3 if (q < 0.81 or 0.91 < q < 0.99) else 1.0) == \
        Q.STATUS_NO_LOCAL_MINIMUM


def test_adaptive_dt_grows_capped_and_matches_fixed():
    fixed_ad, ad = Toy(), Toy()
    mf = _mini(fixed_ad)
    fixed = mf.Get_f_ABC(temp=0.0, meanDensity=0.5, q00=QSTAR)
    m = _mini(ad, dataclasses.replace(RELAX, dtMax=0.5, dtRate=1e-3))
    out = m.Get_f_ABC(temp=0.0, meanDensity=0.5, q00=QSTAR)
    dts = ad.dts
    assert m.last_stop == mf.last_stop == Q.STOP_CONVERGED
    assert dts[0] == pytest.approx(0.1) and 0.15 < max(dts) <= 0.5 + 1e-12
    assert m.last_dt_max == pytest.approx(max(dts))
    assert all(b <= 2.0 * a * (1 + 1e-12) for a, b in zip(dts, dts[1:]))
    assert all(abs(1.0 / d - round(1.0 / d)) < 1e-9 for d in dts)   # each divides the sub-block
    assert ad.from_state == len(dts) - 1
    assert m.last_steps < mf.last_steps
    assert out[0] == pytest.approx(fixed[0], abs=1e-10)


def test_fixed_dt_builds_once():
    ad = Toy()
    m = _mini(ad)
    m.Get_f_ABC(temp=0.0, meanDensity=0.5, q00=QSTAR)
    assert ad.dts == [pytest.approx(0.1)] and ad.builds == 1 and m.last_dt_max == pytest.approx(0.1)
