"""Serving: a continuous-batching loop for PFC relaxation requests.

A request names a model (std or log), a grid side, (r, psibar) and a step
count.  Requests of one shape (model, grid, scheme, dt, precision) run as
slots of a Batch ("slab") with its own HIP stream; each keeps its field and
ETD tables resident in HBM from admission to retirement.  Every tick, each
slab with work advances all of its running requests by the same number of
steps -- the smallest remaining count among them, capped at the slab's tick
length, so every request gets exactly the steps it asked for -- and finished
requests are retired (free energy, mean and optionally the field read back)
while the others keep running.  Slabs step on their own streams, so their
launches overlap on the device.

Tick length and backpressure.  A slab keeps at most two ticks queued on its
stream; the loop skips a slab whose two ticks are still running instead of
waiting for it, so the slabs of small grids advance at their own pace rather
than at the pace of the slowest slab.  A slab with finished requests queues
no new tick until they are retired, and retires them once its stream has
drained, so the read-back does not hold the loop either.  Ticks are `tick_steps` long for grids
of 1024^2 points and more and longer for smaller ones (by the ratio of the
areas, at most 4x), so that small requests need fewer ticks.  (Measured on
the 2000-request stream, this loop: p50 0.24-0.25 s, p99 4.5-4.8 s; with a
blocking wait on every slab's third tick, equal ticks and read-back on the
spot: p50 0.41-0.42 s, p99 5.3-5.5 s; with small-grid ticks 16x longer: p99
7.3-7.6 s.)

Memory admission.  The HBM budget is spent slab by slab: a slab of k slots
reserves k times the per-slot footprint (state and ETD tables, as the
library reports it) when it is created and returns it when it is released.
Admission is first come, first served:

* a request takes a free slot of a slab of its shape that is not draining;
* otherwise a new slab of its shape is created, as large as the shape's
  slabs together (at least `first_slots` and the requests of that shape now
  waiting; at most `max_slots` and what the budget leaves);
* if the budget is spent, idle slabs (no running request) are released,
  least recently used first, until the new slab fits;
* if it still does not fit, the oldest waiting request marks slabs
  "draining" (they admit nothing new) until their footprint covers its
  need; no later request may create a slab before it, so the memory freed
  when the drained slabs empty goes to it;
* a request whose single slot exceeds the whole budget is rejected at once.

Running requests always finish, so draining slabs empty, are released, and
the oldest request is admitted: the loop cannot livelock, and it sleeps
instead of spinning whenever no slab stepped.

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
    rejected: str = ""   # non-empty: the request was refused (the reason)

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


class TorchDevice:
    """The device side of the server: Batches of libpfc_hip on torch streams."""

    def __init__(self):
        import torch
        self.torch = torch

    def slot_bytes(self, params):
        from .plan import slot_bytes
        return slot_bytes(**params)

    def batch(self, capacity, params):
        from .plan import Batch
        return Batch(capacity, **params)

    def stream(self):
        return self.torch.cuda.Stream()

    def using(self, stream):
        return self.torch.cuda.stream(stream)

    def record(self, stream):
        ev = self.torch.cuda.Event()
        ev.record(stream)
        return ev

    def to_host(self, t):
        return t.cpu().numpy()

    def synchronize(self):
        self.torch.cuda.synchronize()


class _Slab:
    """`capacity` slots of one shape on one stream."""

    def __init__(self, dev, key, params, capacity, per_slot, serial):
        self.dev = dev
        self.key = key
        self.params = params
        self.batch = dev.batch(capacity, params)
        self.bytes = int(self.batch.device_bytes())
        self.stream = dev.stream()
        self.capacity = capacity
        self.per_slot = per_slot
        self.serial = serial       # creation order (ties in least-recently-used)
        self.last_use = 0.0        # last admission time
        self.free = deque(range(capacity))
        self.running = {}          # slot -> [request, remaining steps]
        self.flags = np.zeros(capacity, dtype=np.int32)  # sticky non-finite flags per slot
        self.inflight = deque()    # events of enqueued ticks (bounded: backpressure)
        self.dirty = True          # active list needs re-upload
        self.draining = False

    @property
    def idle(self):
        return not self.running

    def admit(self, req, now):
        slot = self.free.popleft()
        self.flags[slot] = 0
        with self.dev.using(self.stream):
            self.batch.set_slot(slot, req.r)
            self.batch.set_field(slot, request_field(req))
        self.running[slot] = [req, req.steps]
        self.dirty = True
        self.last_use = now

    def tick(self, tick_steps):
        """Advance every running request; returns the number of steps taken
        (0 also when two ticks are still queued: the loop comes back)."""
        if not self.running:
            return 0
        while self.inflight and self.inflight[0].query():
            self.inflight.popleft()
        if len(self.inflight) >= 2:
            return 0
        k = min(tick_steps, min(rem for _, rem in self.running.values()))
        if k == 0:  # finished requests wait for retirement: let the stream drain
            return 0
        with self.dev.using(self.stream):
            if self.dirty:
                self.batch.set_active(sorted(self.running))
                self.dirty = False
            self.batch.step(k)
            self.inflight.append(self.dev.record(self.stream))
        for v in self.running.values():
            v[1] -= k
        return k

    def retire(self, t0, keep_fields):
        """Read back and free every finished request."""
        done = [s for s, (_, rem) in self.running.items() if rem == 0]
        if not done:
            return []
        # read back once the slab's queued ticks have run (no new tick is
        # queued meanwhile), so the reductions do not hold the loop
        while self.inflight and self.inflight[0].query():
            self.inflight.popleft()
        if self.inflight:
            return []
        out = []
        with self.dev.using(self.stream):
            self.batch.set_active(done)
            red = self.batch.reduce()
            self.flags |= np.asarray(self.batch.status(), dtype=np.int32)
            fields = [self.dev.to_host(self.batch.get_field(s)) if keep_fields else None for s in done]
        self.inflight.clear()  # the reductions synchronised the stream
        now = time.perf_counter() - t0
        for j, s in enumerate(done):
            req = self.running.pop(s)[0]
            out.append(Result(req.id, float(red[j, 0]), float(red[j, 1]), req.arrival, now, bool(self.flags[s]),
                              fields[j]))
            self.free.append(s)
        self.dirty = True
        return out

    def close(self):
        self.batch.close()


@dataclass
class ServeStats:
    results: list
    wall_s: float
    solves_per_s: float
    p50_s: float
    p99_s: float
    peak_bytes: int
    pools: dict = field(default_factory=dict)
    slabs_created: int = 0
    slabs_released: int = 0
    rejected: int = 0
    slabs_drained: int = 0
    occupied_fraction: float = 0.0   # share of the wall time with at least one request on the device


class Server:
    """Continuous batching over per-shape slabs under an HBM budget."""

    def __init__(self, hbm_budget=200e9, tick_steps=200, precision=32, scheme="etd1", max_slots=64,
                 keep_fields=False, device=None, first_slots=8):
        self.budget = float(hbm_budget)
        self.tick_steps = int(tick_steps)
        self.precision = precision
        self.scheme = scheme
        self.max_slots = int(max_slots)
        self.first_slots = int(first_slots)
        self.keep_fields = keep_fields
        self.dev = device if device is not None else TorchDevice()
        self.slabs = {}        # key -> [slab, ...]
        self.used = 0          # bytes reserved by live slabs
        self.peak = 0
        self.created = 0
        self.released = 0
        self.drained = 0       # slabs marked draining for an older request's memory
        self._per = {}         # key -> bytes per slot
        self._fixed = {}       # key -> bytes of a slab beyond its slots

    def _tick_len(self, slab):
        """Steps per tick of `slab`: tick_steps for grids of >= 1024^2 points,
        longer by the ratio of the areas (at most 4x) below."""
        npts = slab.params["Nx"] * slab.params["Ny"]
        return self.tick_steps * max(1, min(4, (1024 * 1024) // max(1, npts)))

    # -- shapes and memory ----------------------------------------------------
    def _shape(self, req):
        params = request_params(req, self.precision, self.scheme)
        key = (req.model, req.N, self.scheme, params["dt"], self.precision)
        if key not in self._per:
            self._per[key] = int(self.dev.slot_bytes(params))
        return key, params, self._per[key]

    def _all_slabs(self):
        return [s for lst in self.slabs.values() for s in lst]

    def _release(self, slab):
        slab.close()
        self.slabs[slab.key].remove(slab)
        if not self.slabs[slab.key]:
            del self.slabs[slab.key]
        self.used -= slab.bytes
        self.released += 1

    def _release_idle(self, need, keep=None):
        """Release idle slabs, least recently used first, until `need` bytes
        fit in the budget (or none is left)."""
        for slab in sorted((s for s in self._all_slabs() if s.idle and s is not keep),
                           key=lambda s: (s.last_use, s.serial)):
            if self.budget - self.used >= need:
                break
            self._release(slab)
        return self.budget - self.used >= need

    def _need(self, key, per):
        """Bytes a one-slot slab of `key` takes: the slot and the batch's
        fixed part (roots, reduction buffers), measured at the shape's first
        slab; until then one more slot's worth is assumed, or none when no
        slab holds memory (the first slab then measures it)."""
        return per + self._fixed.get(key, per if self.used > 0 else 0)

    def _create(self, key, params, per, nwaiting, now):
        # geometric growth: a shape's next slab holds as many slots as its
        # slabs together (at least `first_slots` and the requests waiting for
        # it), so a steady stream of one shape ends in a few large batches
        have = sum(s.capacity for s in self.slabs.get(key, []))
        want = max(nwaiting, self.first_slots, have)
        fixed = self._need(key, per) - per
        cap = int(max(1, min(self.max_slots, want, (self.budget - self.used - fixed) // per)))
        slab = _Slab(self.dev, key, params, cap, per, self.created)
        self._fixed[key] = max(0, slab.bytes - cap * per)
        slab.last_use = now
        self.created += 1
        self.slabs.setdefault(key, []).append(slab)
        self.used += slab.bytes
        self.peak = max(self.peak, self.used)
        return slab

    def _drain_for(self, need):
        """Mark slabs draining, least recently used first, until the budget
        left plus the footprint of draining and idle slabs covers `need`."""
        slabs = self._all_slabs()
        covered = self.budget - self.used + sum(s.bytes for s in slabs if s.draining or s.idle)
        for slab in sorted((s for s in slabs if not s.draining and not s.idle),
                           key=lambda s: (s.last_use, s.serial)):
            if covered >= need:
                break
            slab.draining = True
            self.drained += 1
            covered += slab.bytes

    def _admit(self, waiting, now, rejected):
        """One first-come-first-served admission pass; returns the requests
        still waiting."""
        # draining slabs that emptied are released at once
        for slab in [s for s in self._all_slabs() if s.draining and s.idle]:
            self._release(slab)
        count = {}
        for req in waiting:
            k = self._shape(req)[0]
            count[k] = count.get(k, 0) + 1
        still = deque()
        blocked = False  # the oldest request that needs memory is waiting for it
        while waiting:
            req = waiting.popleft()
            key, params, per = self._shape(req)
            slab = next((s for s in self.slabs.get(key, []) if s.free and not s.draining), None)
            if slab is None and per + self._fixed.get(key, 0) > self.budget:
                rejected.append(Result(req.id, float("nan"), float("nan"), req.arrival, now, False,
                                       rejected=f"one slot needs {per} bytes, the budget is {int(self.budget)}"))
                count[key] -= 1
                continue
            if slab is None and not blocked:
                need = self._need(key, per)
                if self.budget - self.used < need:
                    self._release_idle(need + (min(count[key], self.max_slots) - 1) * per)
                if self.budget - self.used < need:
                    self._release_idle(need)
                if self.budget - self.used >= need:
                    slab = self._create(key, params, per, count[key], now)
                elif not any(not s.draining for s in self.slabs.get(key, [])):
                    # no slab of its shape will free a slot: claim memory
                    self._drain_for(need)
                    blocked = True
                # else: a running slab of its shape frees a slot when one of
                # its requests finishes; the request keeps its place in line
            if slab is None:
                still.append(req)
                continue
            slab.admit(req, now)
            count[key] -= 1
        return still

    # -- the loop ---------------------------------------------------------------
    def run(self, requests, realtime=True, progress=None, progress_every=10.0):
        """Serve `requests` (sorted by arrival).  With realtime=True a request
        becomes visible at its arrival time; otherwise all are queued at once.
        `progress(done, total, seconds)` is called about every
        `progress_every` seconds."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = deque(reqs)
        results, rejected = [], []
        t0 = time.perf_counter()
        last = t0
        if not realtime:
            for r in reqs:
                r.arrival = 0.0
        waiting = deque()
        occupied = 0.0
        while pending or waiting or any(not s.idle for s in self._all_slabs()):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            if waiting:
                waiting = self._admit(waiting, now, rejected)
            running = any(not s.idle for s in self._all_slabs())
            ticked = 0
            for slab in self._all_slabs():
                ticked += slab.tick(self._tick_len(slab))
            nres = len(results)
            for slab in self._all_slabs():
                results.extend(slab.retire(t0, self.keep_fields))
            if running:
                occupied += time.perf_counter() - t0 - now
            if not running and len(results) == nres:
                if waiting:
                    # nothing runs, so nothing can free memory: admission must
                    # have placed or rejected the oldest request
                    raise RuntimeError("serve: waiting requests with nothing running (admission bug)")
                wait = 0.001
                if pending and not waiting:
                    wait = min(max(0.0, pending[0].arrival - (time.perf_counter() - t0)), 0.01)
                time.sleep(wait)
            elif not ticked and len(results) == nres:
                time.sleep(0.0002)  # every slab has two ticks queued: let them run
            if progress is not None and time.perf_counter() - last >= progress_every:
                last = time.perf_counter()
                progress(len(results), len(reqs), last - t0)
        self.dev.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in results]) if results else np.zeros(1)
        pools = {}
        for s in self._all_slabs():
            pools[str(s.key)] = pools.get(str(s.key), 0) + s.capacity
        return ServeStats(sorted(results + rejected, key=lambda r: r.id), wall, len(results) / wall,
                          float(np.percentile(lat, 50)), float(np.percentile(lat, 99)), self.peak, pools,
                          self.created, self.released, len(rejected), self.drained, occupied / wall)

    def close(self):
        for s in self._all_slabs():
            s.close()
        self.slabs.clear()
        self.used = 0
