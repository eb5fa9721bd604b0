"""The stabilising shift of the logPFC split managed on the device (needs a GPU).

Before every step the library checks the field's range and refits M when it
left the fitted window by more than 10 % (logpfc_extv/simulation.py:332-405),
without a host round trip.  Against the host-driven check (the reference's
per-step ValidatePhiRange / OptimizeM, `stabilise="host"`): the same fits
(M to 1e-13 relative: the device's log and cube may round differently from the
host's in the last place) and the same fields to 1e-11, on a crystallising
instance whose growing range forces refits, on the fast two-kernel path (step
graphs included; the ETD1 step takes the shift through its E1 table and its
row kernel reports the range), on the generic kernels and on the resident
small-grid path.  Batch slots equal
plans bit for bit; a new field refits at the next step.
"""
import numpy as np
import pytest

from pfc_hip import fields as F

pytestmark = pytest.mark.gpu

Q0 = 1.0


def _sim(nx, ny, stabilise, precision=64, seed=5, mean=0.55, scheme="etd1"):
    from pfc_hip.models.logpfc_hip import simulation as S
    geom = S.Build2DGeometry(Nx=nx, Ny=ny, Lx=nx * np.pi / 4, Ly=ny * np.pi / 4)
    params = S.BuildLogPFCModelParameters(Gamma=1.0, beta=0.5, q0=Q0, epsilon=-0.9, g=-2.0, dt=0.1,
                                          meanDensity=mean)
    sim = S.BuildHipGPU_Simulation(params, geom, precision=precision, stabilise=stabilise, scheme=scheme)
    sim.State.phi[...] = F.uniform_noise((ny, nx), mean, 0.05, seed=seed)
    return sim


@pytest.mark.parametrize("shape,generic,scheme", [((96, 64), False, "etd1"), ((96, 64), True, "etd1"),
                                                  ((28, 24), False, "etd1"), ((96, 64), False, "etdrk2")])
def test_device_shift_follows_the_host_check(hip_device, monkeypatch, shape, generic, scheme):
    # (96, 64) ETD1: the fast row kernel reports the range in the step itself,
    # the shift folded into the ETD1 table; generic kernels: the same fold
    # after a separate range pass; (28, 24): the resident stepper; ETDRK2: a
    # range pass before the step, the shift in the row pass
    if generic:
        monkeypatch.setenv("PFC_HIP_GENERIC", "1")
    nx, ny = shape
    dev, host = _sim(nx, ny, True, scheme=scheme), _sim(nx, ny, "host", scheme=scheme)
    assert dev._device_shift and host._host_shift
    fits = set()
    # the range grows from [0.50, 0.60] to ~[0.11, 1.86] in the first ~300
    # steps: short blocks there, then long ones
    for n in [5] * 40 + [300] * 4:
        dev.run(n)
        host.run(n)
        Md, Mh = dev.State.M, host.State.M
        fits.add(round(Mh, 12))
        assert abs(Md - Mh) <= 1e-13 * abs(Mh), (Md, Mh)
        assert abs(dev.State.phi_min - host.State.phi_min) <= 1e-13 * abs(host.State.phi_min)
        assert abs(dev.State.phi_max - host.State.phi_max) <= 1e-13 * abs(host.State.phi_max)
        err = float(np.abs(dev.State.phi.get() - host.State.phi.get()).max())
        assert err <= 1e-11, err
    assert len(fits) >= 3  # the range grew: M was refitted
    assert dev.status() == 0


def test_batch_slots_equal_plans_and_new_field_refits(hip_device):
    from pfc_hip.plan import Batch, Plan
    nx, ny = 64, 48
    kw = dict(model="log", Nx=nx, Ny=ny, Lx=nx * np.pi / 4, Ly=ny * np.pi / 4, beta=0.5, q0=Q0, g=-2.0, dt=0.1,
              clamp_log=True, pin_mean=True)
    fields = [F.uniform_noise((ny, nx), 0.55 + 0.02 * i, 0.05, seed=20 + i) for i in range(3)]
    eps = [-0.9, -0.8, -0.95]
    b = Batch(3, **kw)
    for i in range(3):
        b.set_slot(i, eps[i])
        b.set_field(i, fields[i])
    b.set_auto_shift(True)
    b.step(400)
    for i in range(3):
        pl = Plan(epsilon=eps[i], **kw).set_field(fields[i]).set_auto_shift(True)
        pl.step(400)
        assert np.array_equal(b.get_field(i).cpu().numpy(), pl.get_field().cpu().numpy()), i
        assert b.shift(i) == pl.shift()
    # a new field: no fit until the next step
    pl.set_field(fields[0])
    M, lo, hi = pl.shift()
    assert np.isnan(lo) and np.isnan(hi)
    pl.step(1)
    assert np.isfinite(pl.shift()[1])


def test_device_shift_is_a_logpfc_option(hip_device):
    from pfc_hip import _lib
    from pfc_hip.plan import Plan
    pl = Plan(Nx=64, Ny=64, Lx=50.0, Ly=50.0, epsilon=-0.3, g=0.0, dt=0.5)
    with pytest.raises(_lib.PfcError):
        pl.set_auto_shift(True)
