"""Groups: batches of different grid sizes advanced by shared launches (needs a GPU).

The contract: every slot of every member ends bit-identical to the same batch
stepped on its own with the variant-0 schedule (PFC_HIP_TUNE=0), whatever the
other members are, across graph replays and direct launches, and after an
active list changes.  Members that cannot join (resident-size grids, mixed
precision) are refused with an error.
"""
import numpy as np
import pytest

from pfc_hip import fields as F

pytestmark = pytest.mark.gpu

# (Nx, Ny): square and rectangular, every row and column schedule case of the
# sweep sizes at least once
SHAPES = [(512, 512), (384, 384), (128, 64), (96, 192), (256, 256), (64, 64)]


def _batch(Nx, Ny, nslots, precision, scheme, seed0):
    from pfc_hip.plan import Batch
    Lx, Ly = Nx * np.pi / 4, Ny * np.pi / 4
    b = Batch(nslots, Nx=Nx, Ny=Ny, Lx=Lx, Ly=Ly, g=0.0, dt=0.5, precision=precision, scheme=scheme,
              pin_mean=True)
    for s in range(nslots):
        b.set_slot(s, -0.45 + 0.1 * s)
        b.set_field(s, F.uniform_noise((Ny, Nx), -0.3 + 0.05 * s, 0.05, seed=seed0 + s))
    return b


def _pairs(shapes, nslots, precision, scheme):
    grouped = [_batch(nx, ny, nslots, precision, scheme, 100 * i) for i, (nx, ny) in enumerate(shapes)]
    alone = [_batch(nx, ny, nslots, precision, scheme, 100 * i) for i, (nx, ny) in enumerate(shapes)]
    return grouped, alone


def _same(grouped, alone, tag):
    for i, (g, a) in enumerate(zip(grouped, alone)):
        xg, xa = g.get_fields().cpu().numpy(), a.get_fields().cpu().numpy()
        assert np.array_equal(xg, xa), (tag, i, float(np.abs(xg - xa).max()))
        assert np.array_equal(g.reduce(), a.reduce()), (tag, i)


@pytest.mark.parametrize("precision,scheme", [(32, "etd1"), (64, "etd1"), (64, "etdrk2"), (32, "etdrk2")])
def test_group_bitwise_equal_batches_alone(hip_device, monkeypatch, precision, scheme):
    from pfc_hip.plan import Group
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    grouped, alone = _pairs(SHAPES, 3, precision, scheme)
    grp = Group(grouped)
    for call, n in enumerate([300, 5, 32]):
        grp.step(n)
        for a in alone:
            a.step(n)
        _same(grouped, alone, call)
    assert grp.graph_stats() == (1, 9 + 0 + 1)
    assert not any(b.status().any() for b in grouped)


@pytest.mark.parametrize("precision,scheme", [(32, "etd1"), (64, "etdrk2")])
def test_group_chunks_change_no_result(hip_device, monkeypatch, precision, scheme):
    # a 2 MB cache budget puts the 512^2 and 384^2 members in chunks of their
    # own and the small ones together; each chunk runs its 32 steps in turn
    from pfc_hip.plan import Group
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    monkeypatch.setenv("PFC_HIP_GROUP_CACHE_MB", "2")
    grouped, alone = _pairs(SHAPES, 2, precision, scheme)
    grp = Group(grouped)
    for call, n in enumerate([256, 7, 64]):
        grp.step(n)
        for a in alone:
            a.step(n)
        _same(grouped, alone, call)
    assert grp.chunks() >= 3
    assert grp.graph_stats() == (1, 8 + 2)
    monkeypatch.delenv("PFC_HIP_GROUP_CACHE_MB")
    one = Group(grouped)  # default: one chunk
    one.step(1)
    assert one.chunks() == 1


def test_group_follows_active_lists(hip_device, monkeypatch):
    from pfc_hip.plan import Group
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    shapes = [(256, 128), (64, 64), (192, 192)]
    grouped, alone = _pairs(shapes, 4, 32, "etd1")
    grp = Group(grouped)
    grp.step(300)
    for a in alone:
        a.step(300)
    _same(grouped, alone, "all")
    # one member loses slots, another has none active
    for lst in (grouped, alone):
        lst[0].set_active([3, 1])
        lst[1].set_active([])
    grp.step(50)
    for a in alone:
        a.step(50)
    _same(grouped, alone, "subset")
    for lst in (grouped, alone):
        lst[0].set_active([0, 1, 2, 3])
        lst[1].set_active([2])
    grp.step(260)
    for a in alone:
        a.step(260)
    _same(grouped, alone, "changed again")
    # same counts, other slots and order: the cached graph replays over the
    # lists, which the members update in place
    for lst in (grouped, alone):
        lst[0].set_active([2, 3, 0, 1])
        lst[1].set_active([0])
    grp.step(64)
    for a in alone:
        a.step(64)
    _same(grouped, alone, "same counts")
    assert grp.graph_stats()[0] == 2  # the 50-step call on the subset launched directly


def test_group_refuses_unsupported_members(hip_device):
    from pfc_hip import _lib
    from pfc_hip.plan import Group
    a = _batch(128, 128, 2, 32, "etd1", 1)
    b = _batch(128, 128, 2, 64, "etd1", 1)
    with pytest.raises(_lib.PfcError):
        Group([a, b])  # mixed precision
    small = _batch(16, 16, 2, 32, "etd1", 1)  # resident stepper
    with pytest.raises(_lib.PfcError):
        Group([a, small])
    odd = _batch(160, 160, 1, 32, "etd1", 1)  # 80 and 160 have no group schedule
    with pytest.raises(_lib.PfcError):
        Group([odd])
    with pytest.raises(_lib.PfcError):
        Group([a, a])
