"""Persistent group step (pfc_group_set_flow, k_flow; needs a GPU).

The contract: every slot of every member ends bit-identical to the same batch
stepped on its own with the variant-0 schedule (PFC_HIP_TUNE=0), whatever the
launch length (steps per launch, including a last partial launch), the other
members and the active lists; no slot is flagged aborted.  Members whose row
and column lengths fall in different register classes are refused.
"""
import numpy as np
import pytest

from pfc_hip import fields as F

pytestmark = pytest.mark.gpu

# square and rectangular members covering every class-0 and class-1 length
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
    flowed = [_batch(nx, ny, nslots, precision, scheme, 100 * i) for i, (nx, ny) in enumerate(shapes)]
    alone = [_batch(nx, ny, nslots, precision, scheme, 100 * i) for i, (nx, ny) in enumerate(shapes)]
    return flowed, alone


def _same(flowed, alone, tag):
    for i, (g, a) in enumerate(zip(flowed, alone)):
        xg, xa = g.get_fields().cpu().numpy(), a.get_fields().cpu().numpy()
        assert np.array_equal(xg, xa), (tag, i, float(np.abs(xg - xa).max()))
        assert np.array_equal(g.reduce(), a.reduce()), (tag, i)


@pytest.mark.parametrize("precision,scheme", [(32, "etd1"), (64, "etd1"), (64, "etdrk2"), (32, "etdrk2")])
def test_flow_bitwise_equal_batches_alone(hip_device, monkeypatch, precision, scheme):
    from pfc_hip.plan import Group
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    flowed, alone = _pairs(SHAPES, 3, precision, scheme)
    grp = Group(flowed, flow=16)
    print(f"flow test {precision} {scheme}", flush=True)
    # 40 steps: launches of 16, 16 and 8; then single steps; then one launch
    for call, n in enumerate([40, 1, 16]):
        grp.step(n)
        for a in alone:
            a.step(n)
        _same(flowed, alone, call)
    launches, b0, b1, k = grp.flow_stats()
    assert k == 16 and b0 > 0 and b1 > 0
    assert launches == 2 * (3 + 1 + 1)  # two register classes per launch round
    assert not any(b.status().any() for b in flowed)


def test_flow_one_class_and_active_lists(hip_device, monkeypatch):
    # class-0 members only (one launch chain), slots dropped and re-added
    # between calls: the work list is rebuilt for the new counts
    from pfc_hip.plan import Group
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    shapes = [(256, 256), (128, 64), (64, 64)]
    flowed, alone = _pairs(shapes, 4, 32, "etd1")
    grp = Group(flowed, flow=8)
    grp.step(20)
    for a in alone:
        a.step(20)
    _same(flowed, alone, "all")
    for b in (flowed[0], alone[0]):
        b.set_active([3, 1])
    for b in (flowed[2], alone[2]):
        b.set_active([0])
    grp.step(13)
    for a in alone:
        a.step(13)
    _same(flowed, alone, "subset")
    launches, b0, b1, k = grp.flow_stats()
    assert b1 == 0 and launches == 3 + 2
    assert not any(b.status().any() for b in flowed)


def test_flow_equals_per_pass_group(hip_device, monkeypatch):
    # the two forms of a group step the same tiles: identical fields
    from pfc_hip.plan import Group
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    a = [_batch(nx, ny, 2, 32, "etdrk2", 7 * i) for i, (nx, ny) in enumerate(SHAPES)]
    b = [_batch(nx, ny, 2, 32, "etdrk2", 7 * i) for i, (nx, ny) in enumerate(SHAPES)]
    Group(a, flow=100).step(64)
    Group(b).step(64)
    _same(a, b, "flow vs per-pass")


def test_flow_refuses_mixed_class_member(hip_device):
    from pfc_hip import _lib
    from pfc_hip.plan import Group
    # Nx/2 = 64 (class 0), Ny = 96 (class 1)
    g = Group([_batch(128, 96, 1, 32, "etd1", 0)])
    with pytest.raises(_lib.PfcError):
        g.set_flow(8)
    g.step(2)  # the per-pass form still runs it


def test_flow_abort_flags_every_slot(hip_device, monkeypatch):
    # a dependency wait bounded at zero polls: the first tile that finds its
    # instance's previous phase unfinished stops the launch, every block
    # exits, and every member slot reads PFC_STATUS_ABORTED; the next
    # launch, with the default bound, starts clean
    from pfc_hip.plan import Group
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    monkeypatch.setenv("PFC_HIP_FLOW_SPIN", "0")
    members = [_batch(512, 512, 4, 32, "etd1", 1), _batch(256, 256, 3, 32, "etd1", 9)]
    grp = Group(members, flow=50)
    grp.step(50)
    for b in members:
        assert all(int(f) & 2 for f in b.status()), b.status()
    monkeypatch.delenv("PFC_HIP_FLOW_SPIN")
    alone = [_batch(512, 512, 4, 32, "etd1", 1), _batch(256, 256, 3, 32, "etd1", 9)]
    for i, (nx, ny, n, seed) in enumerate([(512, 512, 4, 1), (256, 256, 3, 9)]):
        for s in range(n):
            members[i].set_field(s, F.uniform_noise((ny, nx), -0.3 + 0.05 * s, 0.05, seed=seed + s))
    grp.step(20)
    for a in alone:
        a.step(20)
    _same(members, alone, "after abort")
    assert not any(b.status().any() for b in members)
