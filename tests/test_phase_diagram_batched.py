"""The batched driver (pfc_hip.phase_diagram.batched) against the serial one
on a toy adapter that implements both the serial names and the batch hooks
with the same arithmetic: every evaluation (values, stop, steps, final
state) and every density-scan result must be identical, rollbacks, NaN stops,
errors and checkpoints included.  The GPU form of this comparison, on the
hip adapters, is in test_gpu_phase_diagram.py."""
import numpy as np
import pytest

from pfc_hip.phase_diagram import batched as B
from pfc_hip.phase_diagram import q0star as Q

QSTAR, CURV, T_BREAK, DT_OK = 1.02, 3.0, 4.0, 0.005
KINDS = {1.5: 'slope', 2.5: 'misfit', 0.7: 'stiff', 0.8: 'blowup', 0.6: 'grow', 0.3: 'raise', -1.0: 'reject'}


class _Sim:
    """A toy instance: (x, y, a) decays (x, y) and, in the 'grow' regime,
    grows the amplitude variable a; everything else is set by (temp, rho,
    q00) and the regime."""

    def __init__(self, temp, rho, q00, dt, state, table):
        self.temp, self.rho, self.q00, self.dt, self.t, self.table = temp, rho, q00, dt, 0.0, table
        self.kind = KINDS.get(float(rho), 'relax')
        self.x, self.y, self.a = (np.array(v, dtype=float) for v in state)

    def step(self):
        self.x *= 1 - 0.4 * self.dt
        self.y *= 1 - 0.9 * self.dt
        if self.kind == 'grow':
            self.a = np.minimum(0.25, self.a * (1 + self.dt))
        self.t += self.dt
        if self.t >= T_BREAK:
            if self.kind == 'blowup' or (self.kind == 'stiff' and self.dt > DT_OK):
                self.x[...] = np.nan
            if self.kind == 'raise':
                raise RuntimeError('toy integrator failed')

    def floor(self):
        if self.table is not None:
            return self.table[0](self.q00)
        if self.kind == 'slope':
            return self.temp + self.q00
        return self.temp + self.rho ** 2 + CURV * (self.q00 - QSTAR) ** 2

    def free_energy(self):
        return self.floor() + np.mean(self.x ** 2) + np.mean(self.y ** 2)

    def diagnose(self, a=None):
        a = self.a if a is None else a
        if self.table is not None:
            fit, amp = self.table[1](self.q00), self.table[2](self.q00)
        else:
            fit = 0.4 if self.kind == 'misfit' else 1.0 - (self.q00 - QSTAR) ** 2
            amp = float(a[0]) if self.kind == 'grow' else 0.25
        return Q.Diagnostics(fitness=fit, amplitudes=(amp, 0.5 * amp))


class ToyBatch:
    def __init__(self, ad, capacity):
        self.ad, self.slots, self.active = ad, [None] * capacity, list(range(capacity))

    def place(self, slot, temp, meanDensity, q00, dt, state, t=0.0):
        self.slots[slot] = self.ad.build({}, 'dots', {}, {}, temp, meanDensity, q00, dt, state=state)
        self.slots[slot].t = t

    def states(self):
        return [self.ad.get_state(self.slots[j]) for j in self.active]

    def set_state(self, slot, state):
        self.ad.set_state(self.slots[slot], state)

    def set_active(self, slots):
        self.active = list(slots)

    def step(self, n):
        for j in self.active:
            for _ in range(n):
                self.slots[j].step()

    def step_each(self, counts):
        for j, n in zip(self.active, counts):
            for _ in range(n):
                self.slots[j].step()
        return True

    def free_energy(self):
        return np.array([self.slots[j].free_energy() for j in self.active])

    def diagnose_active(self, structure, cell):
        return [self.slots[j].diagnose() for j in self.active]

    def close(self):
        self.slots = None


class Toy:
    SLUG = 'toy'
    STRUCTURES = {'dots': 'dots'}

    def __init__(self, table=None, batches=True):
        self.table = table
        self.batches_opened = 0
        if not batches:
            del self.open_batch

    def validate_params(self, p):
        return dict(p)

    validate_cell = validate_ansatz = validate_params

    def ansatz_state(self, structure, cell, ansatz, meanDensity):
        kind = KINDS.get(float(meanDensity), 'relax')
        if self.table is not None or kind == 'grow':
            return (np.zeros(4), np.zeros(6), np.array([2e-6]))
        return (np.linspace(0.2, 0.6, 4) + meanDensity, np.full(6, -0.3), np.array([0.25]))

    def build(self, params, structure, cell, ansatz, temp, meanDensity, q00, dt, state=None, nThreads=1):
        if float(meanDensity) < 0:
            raise ValueError('density out of range')
        if state is None:
            state = self.ansatz_state(structure, cell, ansatz, meanDensity)
        return _Sim(float(temp), float(meanDensity), float(q00), dt, state, self.table)

    def open_batch(self, params, cell, capacity, precision=64):
        self.batches_opened += 1
        return ToyBatch(self, capacity)

    def get_state(self, s):
        return (s.x.copy(), s.y.copy(), s.a.copy())

    def set_state(self, s, state):
        s.x, s.y, s.a = (np.array(v, dtype=float) for v in state)

    def get_time(self, s):
        return s.t

    def set_time(self, s, t):
        s.t = t

    def get_dt(self, s):
        return s.dt

    def step(self, s):
        s.step()

    def free_energy(self, s):
        return s.free_energy()

    def diagnose(self, s, structure, cell):
        return s.diagnose()


RELAX = Q.RelaxSettings(dt=0.1, totalT=2e3, stepT=10.0, divs=10, maxDepth=3, requiredFitness=0.995, minimumA=1e-6)
QSCAN = Q.QScanSettings(passes=3, Deltaq=0.02, q00_Steps=21, Deltaq_Refinement=0.05)


def _mini(ad, relax=RELAX):
    return Q.MinimizeFreeEnergy(ad, {}, 'dots', {}, {}, relax, QSCAN)


def _same(a, b):
    assert (a.f, a.fitness, a.amplitudes, a.stop, a.steps, a.dt_max) == \
        (b.f, b.fitness, b.amplitudes, b.stop, b.steps, b.dt_max) or (np.isnan(a.f) and np.isnan(b.f))
    for u, v in zip(a.state, b.state):
        np.testing.assert_array_equal(u, v)


POINTS = [(0.2, 0.5, 1.0), (0.2, 0.5, 0.9), (0.2, 0.7, 1.0), (0.2, 0.8, 1.0), (0.2, 0.6, 1.0), (0.2, 2.5, 1.0),
          (0.1, 1.5, 1.1), (0.3, 0.45, 1.04)]


def test_batched_evaluations_equal_serial_ones():
    ad = Toy()
    serial = [_mini(ad).evaluate(*p) for p in POINTS]
    batched = B.evaluate_batched(_mini(ad), POINTS)
    assert ad.batches_opened == 1
    for s, b in zip(serial, batched):
        _same(s, b)
    # the regimes reached their stops: converged, poor fit away from q*,
    # rollback then converged, NaN at every dt, the growing seed, poor fit
    assert [e.stop for e in batched[:6]] == [Q.STOP_CONVERGED, Q.STOP_POOR_FIT, Q.STOP_CONVERGED, Q.STOP_NAN,
                                             Q.STOP_CONVERGED, Q.STOP_POOR_FIT]
    assert batched[4].steps == 300


def test_batched_from_given_states_and_without_noise():
    ad = Toy()
    m = _mini(ad, Q.RelaxSettings(**{**RELAX.__dict__, 'noise': 0.0}))
    states = [(np.full(4, 0.1), np.full(6, 0.05), np.array([0.25])), None]
    pts = POINTS[:2]
    serial = [m.evaluate(*p, state=s) for p, s in zip(pts, states)]
    for s, b in zip(serial, B.evaluate_batched(m, pts, states=states)):
        _same(s, b)


def test_batched_find_q0star_and_records_equal_serial():
    r1, r2 = [], []
    q1, f1 = _mini(Toy()).Find_q0Star(0.2, 0.5, record=r1.append)
    q2, f2 = B.find_q0star_batched(_mini(Toy()), 0.2, 0.5, record=r2.append)
    assert (q1, float(f1)) == (q2, float(f2)) and r1 == r2
    assert q1 == pytest.approx(QSTAR, abs=1e-9)


def _table(f, fit, A=lambda q: 0.25):
    return Toy(table=(f, fit, A))


@pytest.mark.parametrize("table", [
    (lambda q: -np.exp(-((q - 1.10) / 0.035) ** 2) - 0.6 * np.exp(-((q - 0.94) / 0.035) ** 2),
     lambda q: 0.3 if q < 0.89 else 1.0),
    (lambda q: 0.5 * (q - 0.96) ** 2 - 0.02 if q < 0.91 else 0.5 * (q - 1.08) ** 2,
     lambda q: 0.3 if 0.91 < q < 0.99 else 1.0),
])
def test_batched_decisions_equal_serial(table):
    def run(fn):
        try:
            return fn()
        except Exception as e:  # noqa: BLE001
            return type(e).__name__
    a = run(lambda: _mini(_table(*table)).Find_q0Star(0.0, 0.5))
    b = run(lambda: B.find_q0star_batched(_mini(_table(*table)), 0.0, 0.5))
    assert a == b if isinstance(a, str) else (a[0], float(a[1])) == (b[0], float(b[1]))


RHOS = [0.5, 1.5, 2.5, -1.0, 0.3, 0.7, 0.8, 0.45]


def _rows(r):
    qs = r['qscan']
    key = np.lexsort((qs['q00'], qs['pass_index'], qs['phi_index']))
    return {k: (v[key] if np.ndim(v) and len(v) == len(key) else v) for k, v in qs.items()}


def _assert_same_scan(a, b):
    for k in ('phi', 'q0star', 'fmin', 'status', 'message'):
        np.testing.assert_array_equal(a[k], b[k])
    ra, rb = _rows(a), _rows(b)
    assert set(ra) == set(rb)
    for k in ra:
        np.testing.assert_array_equal(ra[k], rb[k])


def test_scan_density_batched_equals_serial():
    serial = Q.scan_density(_mini(Toy()), 0.2, RHOS)
    ad = Toy()
    batched = B.scan_density_batched(_mini(ad), 0.2, RHOS)
    _assert_same_scan(serial, batched)
    assert list(batched['status']) == ['ok', 'no_local_minimum', 'insufficient_points', 'error', 'error', 'ok',
                                       'insufficient_points', 'ok']
    # one batch per pass while nothing fails; a failing batch is retried per density
    assert ad.batches_opened >= 3


def test_scan_density_batched_checkpoint_and_resume(tmp_path):
    # no failing batch here: pass 0 finishes two densities, pass 2 is interrupted
    rhos = [0.5, 1.5, 2.5, 0.45]
    fresh = B.scan_density_batched(_mini(Toy()), 0.2, rhos)

    class Interrupt(Toy):
        def open_batch(self, *a, **k):
            if self.batches_opened == 2:
                raise KeyboardInterrupt
            return super().open_batch(*a, **k)
    with pytest.raises(KeyboardInterrupt):
        B.scan_density_batched(_mini(Interrupt()), 0.2, rhos, checkpoint=str(tmp_path))
    with np.load(tmp_path / Q.FMIN_FILE) as z:
        assert [bool(s) for s in z['status']] == [False, True, True, False]
    resumed = B.scan_density_batched(_mini(Toy()), 0.2, rhos, checkpoint=str(tmp_path))
    for k in ('phi', 'q0star', 'fmin', 'status', 'message'):
        np.testing.assert_array_equal(resumed[k], fresh[k])
    # the serial scan resumes from the batched checkpoint files too
    again = Q.scan_density(_mini(Toy()), 0.2, rhos, checkpoint=str(tmp_path))
    for k in ('phi', 'q0star', 'fmin', 'status', 'message'):
        np.testing.assert_array_equal(again[k], fresh[k])


ADAPTIVE = Q.RelaxSettings(**{**RELAX.__dict__, 'dtMax': 0.5, 'dtRate': 1e-3})


def test_adaptive_dt_batched_equals_serial():
    # every slot keeps its own dt: sub-blocks in lockstep with per-slot step
    # counts, dt rebuilds, rollbacks (stiff), NaN stops (blowup), the seed
    ad = Toy()
    serial = [_mini(ad, ADAPTIVE).evaluate(*p) for p in POINTS]
    batched = B.evaluate_batched(_mini(ad, ADAPTIVE), POINTS)
    assert ad.batches_opened == 1
    for s, b in zip(serial, batched):
        _same(s, b)
    assert max(e.dt_max for e in batched) > 0.15          # dt grew somewhere
    assert len({e.steps for e in batched}) > 3


def test_adaptive_dt_falls_back_when_slots_cannot_step_apart():
    class Rigid(Toy):
        def open_batch(self, *a, **k):
            b = super().open_batch(*a, **k)
            b.step_each = lambda counts: False
            return b
    ad = Rigid()
    serial = [_mini(Toy(), ADAPTIVE).evaluate(*p) for p in POINTS[:3]]
    batched = B.evaluate_batched(_mini(ad, ADAPTIVE), POINTS[:3])
    for s, b in zip(serial, batched):
        _same(s, b)


def test_batched_find_q0star_with_adaptive_dt():
    r1, r2 = [], []
    q1, f1 = _mini(Toy(), ADAPTIVE).Find_q0Star(0.2, 0.5, record=r1.append)
    q2, f2 = B.find_q0star_batched(_mini(Toy(), ADAPTIVE), 0.2, 0.5, record=r2.append)
    assert (q1, float(f1)) == (q2, float(f2)) and r1 == r2
