"""Serving: continuous batching keeps each request's result independent of the mix (needs a GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests():
    from pfc_hip.serve import Request
    rng = np.random.default_rng(21)
    out = []
    for i in range(12):
        out.append(Request(i, "std" if i % 3 else "log", int(rng.choice([128, 256])),
                           float(rng.uniform(-0.5, -0.05)), float(rng.uniform(-0.5, -0.1)),
                           int(rng.integers(50, 400)), arrival=0.0, seed=500 + i))
    return out


def test_each_result_same_alone_and_in_the_mix(hip_device, monkeypatch):
    """12 mixed std/log requests through pools of 3 slots (so requests queue,
    retire mid-flight and reuse slots): every result is bit-identical to the
    same request run alone with a Plan (kernel autotuner off so both use the
    same launch shapes)."""
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    from pfc_hip.plan import Plan
    from pfc_hip.serve import Server, request_field, request_params
    reqs = _requests()
    srv = Server(hbm_budget=4e9, tick_steps=64, max_slots=3, keep_fields=True)
    stats = srv.run(reqs, realtime=False)
    srv.close()
    assert [r.id for r in stats.results] == list(range(12))
    for res, req in zip(stats.results, reqs):
        pl = Plan(**request_params(req)).set_field(request_field(req)).step(req.steps)
        fe, mean, _, _ = pl.reduce()
        assert not res.nonfinite
        assert res.fe == fe and res.mean == mean, req
        assert np.array_equal(res.field, pl.get_field().cpu().numpy()), req


def test_poisson_stream_latencies(hip_device):
    """A short real-time Poisson stream: every request completes, time-to-result
    is at least the arrival-to-finish span and p50 <= p99."""
    from pfc_hip.serve import Server, poisson_stream
    reqs = poisson_stream(n=30, rate=200.0, sides=(128, 256), steps=(50, 400))
    srv = Server(hbm_budget=8e9, tick_steps=100, max_slots=8)
    stats = srv.run(reqs, realtime=True)
    srv.close()
    assert len(stats.results) == 30
    assert all(r.latency > 0 for r in stats.results)
    assert 0 < stats.p50_s <= stats.p99_s
    assert stats.solves_per_s > 0


def test_serving_mix_at_its_sizes_under_a_binding_budget(hip_device, monkeypatch):
    """The serving workload's shapes: std and log requests, sides 128 to 2048
    (five 2048^2 requests), log-uniform step counts capped at 2000, FP32
    bulk, arriving as a Poisson stream, under a 0.5 GB budget that cannot hold
    every shape's slab at once (one 8-slot 2048^2 slab takes 0.4 GB; slabs
    are drained and released on the way).
    * every result is bit-identical to the same request run alone with a Plan
      (autotuner off), fields included;
    * one request per (side, model) against the plain FP64 solve: relative
      free energy <= 1e-6;
    * the server never reserved more than the budget."""
    monkeypatch.setenv("PFC_HIP_TUNE", "0")
    from pfc_hip.plan import Plan
    from pfc_hip.serve import Server, poisson_stream, request_field, request_params
    reqs = poisson_stream(n=20, rate=40.0, seed=6, steps=(500, 2000))
    assert sum(q.N == 2048 for q in reqs) == 5 and {q.model for q in reqs} == {"std", "log"}
    budget = 0.5e9
    srv = Server(hbm_budget=budget, tick_steps=200, keep_fields=True)
    print(f"serve mix: 20 requests, budget {budget:.1e}", flush=True)
    stats = srv.run(reqs, realtime=True)
    srv.close()
    print(f"serve mix: wall {stats.wall_s:.2f} s, p50 {stats.p50_s:.2f} s, p99 {stats.p99_s:.2f} s, "
          f"slabs {stats.slabs_created} created / {stats.slabs_released} released / {stats.slabs_drained} drained",
          flush=True)
    assert stats.rejected == 0 and len(stats.results) == 20
    assert stats.peak_bytes <= budget
    assert stats.slabs_released >= 1
    checked = set()
    for res, req in zip(stats.results, reqs):
        assert res.id == req.id and not res.nonfinite
        pl = Plan(**request_params(req)).set_field(request_field(req)).step(req.steps)
        fe, mean, _, _ = pl.reduce()
        assert res.fe == fe and res.mean == mean, req
        assert np.array_equal(res.field, pl.get_field().cpu().numpy()), req
        pl.close()
        if (req.N, req.model) not in checked:
            checked.add((req.N, req.model))
            p64 = Plan(**request_params(req, precision=64)).set_field(request_field(req)).step(req.steps)
            fe64 = p64.reduce()[0]
            p64.close()
            rel = abs(res.fe - fe64) / abs(fe64)
            print(f"serve mix: request {req.id} N={req.N} {req.model} steps={req.steps} FP32 vs FP64 rel FE {rel:.2e}",
                  flush=True)
            assert rel <= 1e-6, (req, rel)
