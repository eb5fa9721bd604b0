"""Serving admission under a binding HBM budget, on a stub device (CPU only).

The stub batch records the steps each slot took and reports them as the
"free energy", so a served result shows that its request got exactly its
steps.  Budgets here hold one to two slabs, so slabs are released and drained
and requests wait for memory; the loop must still serve every request, never
hold more than the budget, and reject a request whose single slot exceeds
the whole budget.
"""
import contextlib

import numpy as np
import pytest

from pfc_hip import serve as SV


class _Event:
    def synchronize(self):
        pass

    def query(self):
        return True


class _Batch:
    def __init__(self, dev, capacity, per):
        self.dev, self.capacity, self.per = dev, capacity, per
        self.done = np.zeros(capacity, dtype=np.int64)
        self.active = []
        self.closed = False
        dev.live += capacity * per
        dev.peak = max(dev.peak, dev.live)

    def device_bytes(self):
        return self.capacity * self.per

    def set_slot(self, slot, eps):
        assert not self.closed
        self.done[slot] = 0

    def set_field(self, slot, field):
        assert field.shape[0] == field.shape[1]

    def set_active(self, slots):
        assert all(0 <= s < self.capacity for s in slots)
        self.active = list(slots)

    def step(self, k):
        assert not self.closed and k >= 0
        for s in self.active:
            self.done[s] += k

    def reduce(self):
        return np.array([[float(self.done[s]), 0.0, 0.0, 0.0] for s in self.active])

    def status(self):
        return np.zeros(self.capacity, dtype=np.int32)

    def get_field(self, slot):
        return np.zeros((2, 2))

    def close(self):
        assert not self.closed
        self.closed = True
        self.dev.live -= self.capacity * self.per


class _Device:
    """Per-slot bytes proportional to the grid area (N^2 bytes)."""

    def __init__(self):
        self.live = 0
        self.peak = 0

    def slot_bytes(self, params):
        return params["Nx"] * params["Ny"]

    def batch(self, capacity, params):
        return _Batch(self, capacity, self.slot_bytes(params))

    def stream(self):
        return object()

    def using(self, stream):
        return contextlib.nullcontext()

    def record(self, stream):
        return _Event()

    def to_host(self, x):
        return x

    def synchronize(self):
        pass


def _requests(n, sides, seed=3, steps=(10, 400)):
    rng = np.random.default_rng(seed)
    return [SV.Request(i, "std" if rng.uniform() < 0.7 else "log", int(rng.choice(sides)), -0.3, -0.3,
                       int(rng.integers(*steps)), 0.0, seed=i) for i in range(n)]


@pytest.mark.parametrize("budget_slots", [1, 3])
def test_binding_budget_serves_everything(budget_slots):
    """Budget = `budget_slots` slots of the largest shape: the server creates,
    drains and releases slabs and serves all 60 requests with exactly their
    steps, never holding more than the budget."""
    dev = _Device()
    reqs = _requests(60, (8, 16, 32))
    budget = budget_slots * 32 * 32
    srv = SV.Server(hbm_budget=budget, tick_steps=50, max_slots=4, device=dev)
    stats = srv.run(reqs, realtime=False)
    srv.close()
    assert stats.rejected == 0
    assert [r.id for r in stats.results] == list(range(60))
    for res, req in zip(stats.results, reqs):
        assert res.fe == req.steps
    assert dev.peak <= budget and stats.peak_bytes <= budget
    assert stats.slabs_released >= 1
    assert dev.live == 0


def test_request_larger_than_the_budget_is_rejected():
    dev = _Device()
    reqs = _requests(20, (8, 16))
    reqs.append(SV.Request(20, "std", 64, -0.3, -0.3, 100, 0.0, seed=20))
    srv = SV.Server(hbm_budget=2 * 16 * 16, tick_steps=50, max_slots=4, device=dev)
    stats = srv.run(reqs, realtime=False)
    srv.close()
    assert stats.rejected == 1
    bad = [r for r in stats.results if r.rejected]
    assert [r.id for r in bad] == [20] and "budget" in bad[0].rejected
    assert all(r.fe == q.steps for r, q in zip(stats.results[:20], reqs[:20]))


def test_oldest_request_is_not_starved_by_its_successors():
    """Eight 8^2 requests fill two 4-slot slabs; a 32^2 request behind them
    needs more memory than is left, and ten more 8^2 requests queue behind
    it.  The 32^2 request drains the least recently used slab, takes the
    freed memory before any later request may create a slab, and finishes
    before every later 8^2 request; the stream drains within the budget."""
    dev = _Device()
    small = [SV.Request(i, "std", 8, -0.3, -0.3, 300, 0.0, seed=i) for i in range(8)]
    big = SV.Request(8, "std", 32, -0.3, -0.3, 50, 0.0, seed=8)
    later = [SV.Request(9 + i, "std", 8, -0.3, -0.3, 300, 0.0, seed=9 + i) for i in range(10)]
    budget = 2 * 4 * 8 * 8 + 32 * 32 - 1  # two 4-slot 8^2 slabs and one byte short of a 32^2 slot
    srv = SV.Server(hbm_budget=budget, tick_steps=20, max_slots=4, first_slots=4, device=dev)
    stats = srv.run(small + [big] + later, realtime=False)
    srv.close()
    assert len(stats.results) == 19 and stats.rejected == 0
    assert stats.slabs_drained >= 1 and stats.slabs_released >= 1
    by_id = {r.id: r for r in stats.results}
    assert by_id[8].fe == 50
    assert all(by_id[i].fe == 300 for i in range(8)) and all(by_id[9 + i].fe == 300 for i in range(10))
    assert by_id[8].done < min(by_id[9 + i].done for i in range(10))
    assert dev.peak <= budget


def test_loop_sleeps_between_sparse_arrivals():
    """Two requests 50 ms apart: the loop waits for the second (no error, no
    busy spin needed for correctness) and serves both."""
    dev = _Device()
    reqs = [SV.Request(0, "std", 8, -0.3, -0.3, 10, 0.0, seed=0), SV.Request(1, "log", 8, -0.3, -0.3, 10, 0.05, seed=1)]
    stats = SV.Server(hbm_budget=1e6, device=dev).run(reqs, realtime=True)
    assert [r.fe for r in stats.results] == [10.0, 10.0]
    assert stats.wall_s >= 0.05
