"""advance_time (CPU): the time a TimeStep loop reaches, to the bit.

The simulations' run(n) and the batched phase-diagram driver advance t by n
steps at once; the serial driver and the reference's callers add dt once per
step.  The two must agree exactly, since the driver's stop tests and records
compare times between the batched and serial forms."""
import numpy as np

from pfc_hip.models._base import advance_time


def _loop(t, dt, n):
    for _ in range(n):
        t += dt
    return t


def test_equals_the_step_loop_bit_for_bit():
    rng = np.random.default_rng(11)
    for _ in range(500):
        t = float(rng.uniform(0.0, 1e3))
        dt = float(rng.choice([0.1, 0.01, 0.5, 1.0 / 3.0, rng.uniform(1e-5, 2.0)]))
        n = int(rng.integers(0, 5000))
        assert advance_time(t, dt, n) == _loop(t, dt, n)


def test_keeps_numpy_scalars():
    t = advance_time(np.float64(0.0), np.float64(0.1), 1000)
    assert isinstance(t, np.float64)
    assert t == _loop(np.float64(0.0), np.float64(0.1), 1000)
    assert t != 100.0  # the loop's rounding, not n * dt
    assert advance_time(2.5, 0.1, 0) == 2.5
