"""Serving: a continuous-batching loop for PFC relaxation requests.

A request names a model (std or log), a grid side, (r, psibar) and a step
count.  Requests of one shape (model, grid, scheme, dt, precision) share a
Batch ("pool") with its own HIP stream; each request occupies one slot,
keeps its field and ETD tables resident in HBM, and is admitted as soon as a
slot is free and the HBM budget allows the pool.  Every tick, each pool with
work advances all of its running requests by the same number of steps -- the
smallest remaining count among them, capped at `tick_steps`, so every request
gets exactly the steps it asked for -- and finished requests are retired
(free energy, mean and optionally the field read back) while the others keep
running.  Pools step on their own streams, so their launches overlap on the
device.

A slot's arithmetic does not depend on its neighbours (the batch contract),
so a request's result is the same alone and in the mix; with the kernel
autotuner disabled (PFC_HIP_TUNE=0) it is bit-identical.

Time-to-result of a request runs from its arrival to the moment its result
is on the host.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import fields as F


@dataclass
class Request:
    id: int
    model: str          # "std" or "log"
    N: int
    r: float
    psibar: float
    steps: int
    arrival: float = 0.0  # seconds after the start of the stream
    seed: int = 0


@dataclass
class Result:
    id: int
    fe: float
    mean: float
    arrival: float
    done: float
    nonfinite: bool
    field: object = None

    @property
    def latency(self):
        return self.done - self.arrival


# model parameters of the served workload: stdPFC at (epsilon = r, psibar);
# logPFC (phi = 1 + psi) at epsilon = r with the builder defaults otherwise
def request_params(req: Request, precision=32, scheme="etd1", dx=math.pi / 4):
    lx, ly = F.commensurate_box(req.N, req.N, dx=dx)
    if req.model == "std":
        return dict(model="std", Nx=req.N, Ny=req.N, Lx=lx, Ly=ly, Gamma=1.0, beta=1.0, q0=1.0, epsilon=req.r,
                    g=0.0, dt=0.5, precision=precision, scheme=scheme, pin_mean=True)
    return dict(model="log", Nx=req.N, Ny=req.N, Lx=lx, Ly=ly, Gamma=1.0, beta=1.0, q0=1.0, epsilon=req.r,
                g=1.0, dt=0.1, minPhi=1e-6, precision=precision, scheme=scheme, clamp_log=True, pin_mean=True)


def request_field(req: Request):
    if req.model == "std":
        return F.uniform_noise((req.N, req.N), req.psibar, 0.05, seed=req.seed)
    return F.log_noise((req.N, req.N), psibar=req.psibar, half_width=0.05, seed=req.seed)


def poisson_stream(n=2000, rate=100.0, seed=5, sides=(128, 256, 512, 1024, 2048), steps=(500, 20000),
                   std_fraction=0.7, r_range=(-0.5, -0.05), psi_range=(-0.6, 0.0)):
    """Seeded requests: Poisson arrivals at `rate` per second, 70 % stdPFC,
    grid sides log-uniform over powers of two, step counts log-uniform."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        model = "std" if rng.uniform() < std_fraction else "log"
        N = int(rng.choice(sides))
        nst = int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))))
        r = float(rng.uniform(*r_range))
        pb = float(rng.uniform(*psi_range))
        out.append(Request(i, model, N, r, pb, nst, t, seed=10_000 + i))
    return out


class _Pool:
    def __init__(self, key, params, capacity):
        import torch
        from .plan import Batch
        self.key = key
        self.params = params
        self.batch = Batch(capacity, **params)
        self.stream = torch.cuda.Stream()
        self.capacity = capacity
        self.free = deque(range(capacity))
        self.running = {}      # slot -> [request, remaining steps]
        self.flags = np.zeros(capacity, dtype=np.int32)  # sticky non-finite flags per slot
        self.inflight = deque()  # events of enqueued ticks (bounded: backpressure)
        self.dirty = True      # active list needs re-upload

    def admit(self, req, now):
        import torch
        slot = self.free.popleft()
        self.flags[slot] = 0
        with torch.cuda.stream(self.stream):
            self.batch.set_slot(slot, req.r)
            self.batch.set_field(slot, request_field(req))
        self.running[slot] = [req, req.steps]
        self.dirty = True

    def tick(self, tick_steps):
        """Advance every running request; returns the number of steps taken."""
        import torch
        if not self.running:
            return 0
        k = min(tick_steps, min(rem for _, rem in self.running.values()))
        with torch.cuda.stream(self.stream):
            if self.dirty:
                self.batch.set_active(sorted(self.running))
                self.dirty = False
            self.batch.step(k)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.inflight.append(ev)
        # keep at most two ticks queued per pool so newly admitted requests
        # do not wait behind a long backlog
        while len(self.inflight) > 2:
            self.inflight.popleft().synchronize()
        for v in self.running.values():
            v[1] -= k
        return k

    def retire(self, t0, keep_fields):
        """Read back and free every finished request."""
        import torch
        done = [s for s, (_, rem) in self.running.items() if rem == 0]
        if not done:
            return []
        out = []
        with torch.cuda.stream(self.stream):
            self.batch.set_active(done)
            red = self.batch.reduce()
            self.flags |= self.batch.status()
            fields = [self.batch.get_field(s).cpu().numpy() if keep_fields else None for s in done]
        self.inflight.clear()  # the reductions synchronised the stream
        now = time.perf_counter() - t0
        for j, s in enumerate(done):
            req = self.running.pop(s)[0]
            out.append(Result(req.id, float(red[j, 0]), float(red[j, 1]), req.arrival, now, bool(self.flags[s]),
                              fields[j]))
            self.free.append(s)
        self.dirty = True
        return out


@dataclass
class ServeStats:
    results: list
    wall_s: float
    solves_per_s: float
    p50_s: float
    p99_s: float
    peak_bytes: int
    pools: dict = field(default_factory=dict)


class Server:
    """Continuous batching over per-shape pools under an HBM budget."""

    def __init__(self, hbm_budget=200e9, tick_steps=200, precision=32, scheme="etd1", max_slots=64,
                 keep_fields=False):
        self.budget = float(hbm_budget)
        self.tick_steps = int(tick_steps)
        self.precision = precision
        self.scheme = scheme
        self.max_slots = max_slots
        self.keep_fields = keep_fields
        self.pools = {}
        self.used = 0

    def _pool_for(self, req):
        from .plan import slot_bytes
        params = request_params(req, self.precision, self.scheme)
        key = (req.model, req.N, self.scheme, params["dt"], self.precision)
        pool = self.pools.get(key)
        if pool is None:
            per = slot_bytes(**params)
            cap = int(max(1, min(self.max_slots, (self.budget - self.used) // (2 * per))))
            if per * cap > self.budget - self.used:
                return None, key
            pool = _Pool(key, params, cap)
            self.pools[key] = pool
            self.used += pool.batch.device_bytes()
        return pool, key

    def run(self, requests, realtime=True, progress=None, progress_every=10.0):
        """Serve `requests` (sorted by arrival).  With realtime=True a request
        becomes visible at its arrival time; otherwise all are queued at once.
        `progress(done, total, seconds)` is called about every
        `progress_every` seconds."""
        import torch
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = deque(reqs)
        results = []
        t0 = time.perf_counter()
        last = t0
        if not realtime:
            for r in reqs:
                r.arrival = 0.0
        waiting = deque()
        while pending or waiting or any(p.running for p in self.pools.values()):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: FIFO, a request waits only for its own pool's slots
            still = deque()
            while waiting:
                req = waiting.popleft()
                pool, _ = self._pool_for(req)
                if pool is not None and pool.free:
                    pool.admit(req, now)
                else:
                    still.append(req)
            waiting = still
            stepped = 0
            for pool in self.pools.values():
                stepped += pool.tick(self.tick_steps)
            for pool in self.pools.values():
                results.extend(pool.retire(t0, self.keep_fields))
            if progress is not None and time.perf_counter() - last >= progress_every:
                last = time.perf_counter()
                progress(len(results), len(reqs), last - t0)
            if stepped == 0 and pending and not waiting:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in results]) if results else np.zeros(1)
        return ServeStats(sorted(results, key=lambda r: r.id), wall, len(results) / wall,
                          float(np.percentile(lat, 50)), float(np.percentile(lat, 99)), self.used,
                          {str(k): p.capacity for k, p in self.pools.items()})

    def close(self):
        for p in self.pools.values():
            p.batch.close()
