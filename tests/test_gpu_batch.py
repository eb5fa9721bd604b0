"""Batches: many same-shape instances per launch (needs a GPU).

The contract: a slot's result does not depend on what else is in the batch.
With the kernel schedule pinned it is bit-identical to the same instance run
alone; with the autotuner free to pick shapes per batch size it agrees to
1e-13 (FP64).  A non-finite value in one slot is flagged for that slot only.
"""
import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

pytestmark = pytest.mark.gpu

N = 128
L = N * np.pi / 4
EPS = [-0.45, -0.3, -0.2, -0.1, -0.05, -0.25]
PSIBAR = [-0.3, -0.25, -0.2, -0.35, -0.15, 0.0]


def _kw(precision=64, scheme="etdrk2"):
    return dict(Nx=N, Ny=N, Lx=L, Ly=L, g=0.0, dt=0.5, precision=precision, scheme=scheme)


def _fields():
    return [F.uniform_noise((N, N), pb, 0.05, seed=100 + i) for i, pb in enumerate(PSIBAR)]


def _batch(fields, precision=64, scheme="etdrk2"):
    from pfc_hip.plan import Batch
    b = Batch(len(fields), **_kw(precision, scheme))
    for i, f in enumerate(fields):
        b.set_slot(i, EPS[i])
        b.set_field(i, f)
    return b


def _alone(i, f, precision=64, scheme="etdrk2"):
    from pfc_hip.plan import Plan
    return Plan(epsilon=EPS[i], **_kw(precision, scheme)).set_field(f)


def test_slot_bitwise_equal_alone_with_pinned_schedule(hip_device, monkeypatch):
    monkeypatch.setenv("PFC_HIP_SCHEDULE", "0:8:0:4")
    fields = _fields()
    b = _batch(fields).step(30)
    for i, f in enumerate(fields):
        a = _alone(i, f).step(30)
        assert np.array_equal(b.get_field(i).cpu().numpy(), a.get_field().cpu().numpy()), i


@pytest.mark.parametrize("precision", [64, 32])
def test_slots_match_alone_autotuned(hip_device, precision):
    fields = _fields()
    b = _batch(fields, precision).step(30)
    red = b.reduce()
    tol = 1e-13 if precision == 64 else 2e-5
    for i, f in enumerate(fields):
        a = _alone(i, f, precision).step(30)
        assert np.abs(b.get_field(i).cpu().numpy() - a.get_field().cpu().numpy()).max() <= tol, i
        fe, mean, mn, mx = a.reduce()
        assert red[i][0] == pytest.approx(fe, rel=1e-12 if precision == 64 else 1e-6)
        assert red[i][1] == pytest.approx(mean, abs=1e-15 if precision == 64 else 1e-7)


def test_slots_match_oracle(hip_device):
    fields = _fields()
    b = _batch(fields, scheme="etd1").step(50)
    geom = O.build_geometry(N, N, L, L)
    for i in (0, 3, 5):
        sim = O.Simulation(O.Params(model="std", epsilon=EPS[i], g=0.0, dt=0.5), geom)
        sim.set_field(fields[i])
        sim.step(50)
        assert np.abs(b.get_field(i).cpu().numpy() - sim.field).max() <= 1e-10, i


def test_active_subset_leaves_other_slots_untouched(hip_device):
    fields = _fields()
    b = _batch(fields)
    before = [b.get_field(i).cpu().numpy() for i in range(len(fields))]
    b.set_active([4, 1, 3])
    b.step(10)
    red = b.reduce()
    assert red.shape == (3, 4)
    for i in range(len(fields)):
        now = b.get_field(i).cpu().numpy()
        if i in (1, 3, 4):
            a = _alone(i, fields[i]).step(10)
            assert np.abs(now - a.get_field().cpu().numpy()).max() <= 1e-13
        else:
            assert np.array_equal(now, before[i])
    # reduce order follows the active list
    for j, i in enumerate([4, 1, 3]):
        assert red[j][1] == pytest.approx(float(np.mean(fields[i])), abs=1e-13)


def test_nonfinite_slot_is_isolated(hip_device):
    """Failure handling forced from the host: a NaN uploaded into one slot is
    flagged for that slot only and leaves every other slot's trajectory as it
    would be without it."""
    fields = _fields()
    bad = fields[2].copy()
    bad[5, 7] = np.nan
    clean = _batch(fields).step(5)
    b = _batch([*fields[:2], bad, *fields[3:]]).step(5)
    flags = b.status()
    assert flags[2] != 0 and all(flags[i] == 0 for i in range(len(fields)) if i != 2)
    for i in range(len(fields)):
        if i != 2:
            assert np.array_equal(b.get_field(i).cpu().numpy(), clean.get_field(i).cpu().numpy()), i
    assert not np.isfinite(b.get_field(2).cpu().numpy()).all()


def test_set_slot_rebuilds_tables(hip_device):
    fields = _fields()
    b = _batch(fields)
    b.set_slot(0, -0.33)
    b.set_field(0, fields[0])
    b.set_active([0]).step(20)
    from pfc_hip.plan import Plan
    a = Plan(epsilon=-0.33, **_kw()).set_field(fields[0]).step(20)
    assert np.abs(b.get_field(0).cpu().numpy() - a.get_field().cpu().numpy()).max() <= 1e-13


def test_slot_bytes_and_mixed_sizes(hip_device):
    from pfc_hip.plan import Batch, slot_bytes
    per = slot_bytes(**_kw())
    # FP64 state is stored in tiles of 2 kx x 4 rows: kx padded to even;
    # ETD tables keep rows ky <= N/2 (they depend on ky^2 only)
    nc = N * (2 * ((N // 2 + 2) // 2))
    nt = (N // 2 + 1) * (N // 2 + 1)
    grid = (N // 2 + 1) + N + 1  # the slot's kx^2, ky^2 and dt
    assert per == nc * 16 * 3 + nt * 8 * 3 + grid * 8  # state (S, T, D) + FP64 tables (E1, Q1, Q2) + grid
    b = Batch(3, **_kw())
    assert b.device_bytes() >= 3 * per
    # FP32 bulk: the same with the FP32 copies of the three tables counted,
    # and the batch holds at least three such slots
    per32 = slot_bytes(**_kw(precision=32))
    st32 = per32 - nt * 12 * 3 - grid * 8
    assert st32 > 0 and st32 % (8 * 3) == 0  # state (S, T, D) in FP32 complex
    assert Batch(3, **_kw(precision=32)).device_bytes() >= 3 * per32
