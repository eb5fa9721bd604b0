"""Step-loop graphs (needs a GPU).

A step call of n >= 32 steps on the two-kernel path replays the captured HIP
graph of 32 steps of its launch configuration n // 32 times and launches the
rest directly; a configuration gets its graph from the first call of n >= 256
steps (shorter calls on a new configuration launch directly).  The graph's
nodes are the direct launches, so a plan stepped through graphs must equal,
bit for bit, the same plan stepped in calls of fewer than 32 steps (no graph).
Each comparison follows several replays with eager work and synchronisations
in between (free energy, field reads), and the graph must survive changes of
what it reads from device memory (tables after a new epsilon) and be replaced
when a launch dimension changes (the active count).
"""
import numpy as np
import pytest

from pfc_hip import fields as F

pytestmark = pytest.mark.gpu


def _eager(obj, n):
    while n > 0:
        k = min(n, 31)
        obj.step(k)
        n -= k
    return obj


@pytest.mark.parametrize("precision,scheme,n", [(64, "etdrk2", 256), (32, "etd1", 256), (64, "etd1", 128)])
def test_plan_graph_bitwise_equal_eager(hip_device, precision, scheme, n):
    from pfc_hip.plan import Plan
    L = n * np.pi / 4
    kw = dict(Nx=n, Ny=n, Lx=L, Ly=L, epsilon=-0.3, g=0.0, dt=0.5, precision=precision, scheme=scheme)
    psi0 = F.uniform_noise((n, n), -0.3, 0.05, seed=11)
    a = Plan(**kw).set_field(psi0)
    b = Plan(**kw).set_field(psi0)
    for call, nsteps in enumerate([100, 300, 100, 32, 45]):
        a.step(nsteps)
        _eager(b, nsteps)
        fa, fb = a.free_energy(), b.free_energy()
        xa, xb = a.get_field().cpu().numpy(), b.get_field().cpu().numpy()
        assert np.array_equal(xa, xb), (call, float(np.abs(xa - xb).max()))
        assert fa == fb, call
    cap, rep = a.graph_stats()
    assert cap == 1 and rep == 0 + 9 + 3 + 1 + 1, (cap, rep)
    assert b.graph_stats() == (0, 0)


def test_batch_graph_follows_tables_and_active_list(hip_device):
    from pfc_hip.plan import Batch
    n = 128
    L = n * np.pi / 4
    kw = dict(Nx=n, Ny=n, Lx=L, Ly=L, g=0.0, dt=0.5, precision=32, scheme="etdrk2")
    eps = [-0.4, -0.3, -0.2, -0.1]
    fields = [F.uniform_noise((n, n), -0.25, 0.05, seed=20 + i) for i in range(4)]
    bs = []
    for _ in range(2):
        b = Batch(4, **kw)
        for i, f in enumerate(fields):
            b.set_slot(i, eps[i])
            b.set_field(i, f)
        bs.append(b)
    a, e = bs

    def same(tag):
        ra, re_ = a.reduce(), e.reduce()
        xa = a.get_fields().cpu().numpy()
        xe = e.get_fields().cpu().numpy()
        assert np.array_equal(xa, xe), tag
        assert np.array_equal(ra, re_), tag

    a.step(288)
    _eager(e, 288)
    same("all slots")
    # new tables for slot 2: the same graph reads them
    a.set_slot(2, -0.35)
    e.set_slot(2, -0.35)
    a.step(64)
    _eager(e, 64)
    same("new epsilon")
    assert a.graph_stats()[0] == 1
    # a shorter active list changes the launch grid: a short call launches
    # directly, a long one captures a second graph
    a.set_active([3, 1])
    e.set_active([3, 1])
    a.step(70)
    _eager(e, 70)
    same("two active, short call")
    assert a.graph_stats() == (1, 9 + 2)
    a.step(300)
    _eager(e, 300)
    same("two active")
    a.set_active([0, 1, 2, 3])
    e.set_active([0, 1, 2, 3])
    a.step(32)
    _eager(e, 32)
    same("all again")
    cap, rep = a.graph_stats()
    assert cap == 2 and rep == 9 + 2 + 9 + 1, (cap, rep)
