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
plans bit for bit; a new field refits at the next step.  On the fast ETD1
path the decision happens inside the step's two launches (breach flags by
step parity): graphs captured at either parity and batch slots rejoining the
active list after a refit step as plans do.
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


def _log_kw(nx, ny):
    return dict(model="log", Nx=nx, Ny=ny, Lx=nx * np.pi / 4, Ly=ny * np.pi / 4, beta=0.5, q0=Q0, g=-2.0, dt=0.1,
                clamp_log=True, pin_mean=True)


def test_in_step_decision_graphs_at_either_parity(hip_device):
    """The fast ETD1 step decides refits inside its two launches, the row
    tiles raising breach flags that alternate with the step's parity.  Step
    graphs captured after an odd number of steps (a 5-step call, then 300 and
    33) give the same fields and fits, bit for bit, as one 338-step call
    (graphs captured at even parity), through the refits of a crystallising
    field."""
    from pfc_hip.plan import Plan
    kw = dict(_log_kw(96, 64), epsilon=-0.9)
    f0 = F.uniform_noise((64, 96), 0.55, 0.05, seed=5)
    a = Plan(**kw).set_field(f0).set_auto_shift(True)
    b = Plan(**kw).set_field(f0).set_auto_shift(True)
    a.step(5)
    M5 = a.shift()[0]
    a.step(300)
    a.step(33)
    b.step(338)
    assert a.shift()[0] != M5  # refitted inside the graphs
    assert np.array_equal(a.get_field().cpu().numpy(), b.get_field().cpu().numpy())
    assert a.shift() == b.shift()


def test_batch_slot_rejoining_after_a_refit_matches_its_plan(hip_device):
    """A batch slot that refits at its last step before leaving the active
    list, sits out one step (so it rejoins at the parity of that refit's
    flag) and is stepped on: its field and fit equal a plan stepped the same
    number of steps, bit for bit (no stale flag, no spurious refit)."""
    from pfc_hip.plan import Batch, Plan
    kw = _log_kw(64, 48)
    fields = [F.uniform_noise((48, 64), 0.55 + 0.02 * i, 0.05, seed=20 + i) for i in range(3)]
    eps = [-0.9, -0.8, -0.95]
    probe = Plan(epsilon=eps[1], **kw).set_field(fields[1]).set_auto_shift(True)
    Ms = []
    for _ in range(120):
        probe.step(1)
        Ms.append(probe.shift()[0])
    # k steps end with a refit, and step k + 1 keeps the fit
    k = next(i + 1 for i in range(2, len(Ms) - 1) if Ms[i] != Ms[i - 1] and Ms[i + 1] == Ms[i])
    b = Batch(3, **kw)
    for i in range(3):
        b.set_slot(i, eps[i])
        b.set_field(i, fields[i])
    b.set_auto_shift(True)
    b.step(k)
    b.set_active([0, 2])
    b.step(1)
    b.set_active([0, 1, 2])
    b.step(300)
    for i, n in ((0, k + 301), (1, k + 300), (2, k + 301)):
        pl = Plan(epsilon=eps[i], **kw).set_field(fields[i]).set_auto_shift(True)
        pl.step(n)
        assert np.array_equal(b.get_field(i).cpu().numpy(), pl.get_field().cpu().numpy()), i
        assert b.shift(i) == pl.shift(), i


def test_device_shift_is_a_logpfc_option(hip_device):
    from pfc_hip import _lib
    from pfc_hip.plan import Plan
    pl = Plan(Nx=64, Ny=64, Lx=50.0, Ly=50.0, epsilon=-0.3, g=0.0, dt=0.5)
    with pytest.raises(_lib.PfcError):
        pl.set_auto_shift(True)


@pytest.mark.parametrize("shape", [(96, 64), (28, 24)])
def test_chemical_potential_is_shift_independent(hip_device, shape):
    """mu = L_mu phi_hat + N_hat does not depend on the stabilising shift (+M in
    L_mu, -M phi in N cancel).  After the device refitted M on a crystallising
    instance, the chemical potential equals the oracle's unshifted one of the
    same field (the filter reads the device's M, as the step does)."""
    from oracle import pfc_oracle as O
    nx, ny = shape
    sim = _sim(nx, ny, True)
    sim.run(200)
    M = sim.State.M
    assert M != 0.0 and abs(M) > 1e-3
    phi = sim.State.phi.get()
    sim._Batch1_CalcNonlinMu()
    sim._Batch2_CalcMu_Terms()
    g = sim.Geometry
    ref = O.Simulation(O.Params(model="log", beta=0.5, epsilon=-0.9, g=-2.0, dt=0.1, clamp_log=True),
                       O.build_geometry(nx, ny, g.Lx, g.Ly))
    ref.set_field(phi)
    mu, lmu = ref.chemical_potential()
    scale = float(np.abs(mu).max())
    assert float(np.abs(sim.State.mu - mu).max()) <= 1e-12 * scale
    assert float(np.abs(sim.State.local_mu - lmu).max()) <= 1e-12 * scale


def test_tolerance_change_keeps_the_fit_and_reaches_the_step(hip_device):
    """State.phi_floor_tol / phi_ceiling_tol are read by the device's check from
    the next step on, without a refit; stepping through step graphs (calls of
    >= 256 steps) picks up the new values as well.  Against the host check
    with the same tolerance change."""
    dev, host = _sim(96, 64, True), _sim(96, 64, "host")
    for s in (dev, host):
        s.run(300)
    M0 = dev.State.M
    for s in (dev, host):
        s.State.phi_floor_tol = 0.02
        s.State.phi_ceiling_tol = 0.02
    assert dev.State.M == M0  # no refit from the change itself
    for n in (5, 5, 300, 300):
        dev.run(n)
        host.run(n)
        assert abs(dev.State.M - host.State.M) <= 1e-13 * abs(host.State.M)
    err = float(np.abs(dev.State.phi.get() - host.State.phi.get()).max())
    assert err <= 1e-11, err


def test_extv_chemical_potential_returns_mu_only(hip_device):
    """logpfc_extv's CalculateChemicalPotential returns mu as one host array
    (models/logpfc_extv/simulation.py:419-425)."""
    from pfc_hip.models.logpfc_extv_hip import simulation as X
    geom = X.BuildGeometry(64, 48, 64 * np.pi / 4, 48 * np.pi / 4)
    params = X.BuildLogPFCModel_withVacancyForce_Parameters(beta=0.5, epsilon=-0.9, g=-2.0, dt=0.1,
                                                            meanDensity=0.55)
    sim = X.BuildHipGPU_Simulation(params, geom)
    sim.State.phi[...] = F.uniform_noise((48, 64), 0.55, 0.05, seed=9)
    sim.run(50)
    mu = sim.CalculateChemicalPotential()
    assert isinstance(mu, np.ndarray) and mu.shape == (48, 64)
    assert np.array_equal(mu, sim.State.mu)


def test_fp32_device_shift_follows_the_host_check(hip_device):
    # FP32 bulk on the fast ETD1 path: the column kernel's in-step refit
    # writes the FP32 table copies the FP32 combine reads (as k_tables does
    # for the host-driven rebuild); the two runs refit at the same steps to
    # close fits and stay within FP32 rounding of each other
    dev, host = _sim(96, 64, True, precision=32), _sim(96, 64, "host", precision=32)
    assert dev._device_shift and host._host_shift
    fits = set()
    for n in [5] * 40 + [300] * 4:
        dev.run(n)
        host.run(n)
        Md, Mh = dev.State.M, host.State.M
        fits.add(round(Mh, 6))
        assert abs(Md - Mh) <= 1e-5 * abs(Mh), (Md, Mh)
        err = float(np.abs(dev.State.phi.get() - host.State.phi.get()).max())
        assert err <= 1e-4, err
    assert len(fits) >= 3
    assert dev.status() == 0
