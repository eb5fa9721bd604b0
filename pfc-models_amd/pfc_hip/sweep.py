"""Phase-diagram sweeps: many stdPFC relaxations at once.

The workload is a 16 x 16 grid of (r, psibar) state points, r in
[-0.5, -0.05] and psibar in [-0.6, 0], each relaxed from psibar +
U(-0.05, 0.05) for a fixed number of ETD1 steps on its own grid.  Grid sides
are drawn from {64, 128, 192, 256, 384, 512} (seeded), and each box holds an
integer number of triangular periods (`fields.commensurate_box`).

Two ways to run it:

* `run_plain`   one instance at a time, one Plan each, FP64 (the plain
                one-job-at-a-time GPU step);
* `run_batched` instances grouped by grid size into Batches (FP32 bulk,
                FP64 tables and reductions); every group advances with one
                row and one column launch per step, groups on their own HIP
                streams so their launches overlap.

Both return per-instance FP64 free energy per site and mean density, which is
what the common-tangent construction consumes (`(phi, fmin)` tables, cf.
phase_diagram/common_tangent.py:336-403 in the reference); `coexistence_by_r`
applies it (`coexistence.py`) at every r of the sweep.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import fields as F

SIZES = (64, 128, 192, 256, 384, 512)


@dataclass(frozen=True)
class Instance:
    index: int
    r: float
    psibar: float
    N: int
    Lx: float
    Ly: float
    seed: int


def sweep_instances(n_r=16, n_psi=16, r_range=(-0.5, -0.05), psi_range=(-0.6, 0.0), size_seed=4,
                    sizes=SIZES, dx=math.pi / 4):
    """The (r, psibar, N) grid; instance i uses noise seed 1000 + i."""
    rng = np.random.default_rng(size_seed)
    rs = np.linspace(*r_range, n_r)
    ps = np.linspace(*psi_range, n_psi)
    out = []
    k = 0
    for r in rs:
        for pb in ps:
            n = int(rng.choice(sizes))
            lx, ly = F.commensurate_box(n, n, dx=dx)
            out.append(Instance(k, float(r), float(pb), n, lx, ly, 1000 + k))
            k += 1
    return out


def initial_field(inst: Instance):
    return F.uniform_noise((inst.N, inst.N), inst.psibar, 0.05, seed=inst.seed)


def _model_kw(inst: Instance, dt, precision, scheme):
    # F = integral psi/2 (r + (1 + lap)^2) psi + psi^4/4: beta = q0 = Gamma = 1, epsilon = r, g = 0
    return dict(model="std", Nx=inst.N, Ny=inst.N, Lx=inst.Lx, Ly=inst.Ly, Gamma=1.0, beta=1.0, q0=1.0,
                epsilon=inst.r, g=0.0, dt=dt, precision=precision, scheme=scheme, pin_mean=True)


@dataclass
class SweepResult:
    fe: np.ndarray        # free energy per site (FP64), by instance index
    mean: np.ndarray      # mean density (FP64)
    wall_s: float
    solves_per_s: float
    nonfinite: np.ndarray  # status flags
    fe_prev: np.ndarray = None  # free energy `checkpoint` steps before the end (run_plain)


def run_plain(instances, steps, dt=0.5, precision=64, scheme="etd1", checkpoint=None):
    """One instance at a time, one Plan each.  With `checkpoint`, the free
    energy is also read `checkpoint` steps before the end (how far each
    relaxation still moves)."""
    import torch
    from .plan import Plan
    n = len(instances)
    fe, mean, bad = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
    prev = np.zeros(n) if checkpoint else None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j, inst in enumerate(instances):
        pl = Plan(**_model_kw(inst, dt, precision, scheme))
        pl.set_field(initial_field(inst))
        if checkpoint:
            pl.step(steps - checkpoint)
            prev[j] = pl.reduce()[0]
            pl.step(checkpoint)
        else:
            pl.step(steps)
        fe[j], mean[j], _, _ = pl.reduce()
        bad[j] = pl.status()
        pl.close()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return SweepResult(fe, mean, wall, n / wall, bad, prev)


def rounding_floor(cases, steps, seed=0):
    """How far the FP64 solve itself moves under perturbations at the level of
    rounding: for each (plan keyword arguments, initial field) in `cases`,
    `steps` FP64 steps from the field, from the field rounded to FP32 (what an
    FP32 state holds) and from the field with every value moved by one ulp
    (random signs).  Returns (fe64, rel_fp32_ic, rel_ulp): the FP64 free
    energies and the two relative free-energy changes.  A relaxation that
    amplifies round-off (nucleation under way) shows it here: an FP32 result
    cannot be held to a tighter bound than the FP32-rounded initial field
    gives the FP64 solve."""
    from .plan import Plan
    rng = np.random.default_rng(seed)
    fe = np.zeros(len(cases))
    ic = np.zeros(len(cases))
    ulp = np.zeros(len(cases))
    for j, (kw, f0) in enumerate(cases):
        f0 = np.asarray(f0, dtype=np.float64)
        variants = (f0, f0.astype(np.float32).astype(np.float64),
                    f0 + np.spacing(f0) * rng.choice([-1.0, 1.0], size=f0.shape))
        out = []
        for v in variants:
            pl = Plan(**{**kw, "precision": 64}).set_field(v).step(steps)
            out.append(pl.reduce()[0])
            pl.close()
        fe[j] = out[0]
        ic[j] = abs(out[1] - out[0]) / abs(out[0])
        ulp[j] = abs(out[2] - out[0]) / abs(out[0])
    return fe, ic, ulp


def fp32_spread(cases, steps, trials=3, seed=0):
    """The FP32 path's own spread: for each (plan keyword arguments, initial
    field), the relative free-energy error against FP64 of the FP32 solve from
    the field and from `trials` copies moved by up to one FP32 ulp (random
    -1/0/+1 per value).  Returns an (n, trials + 1) array.  Where a
    relaxation amplifies round-off, this range is how precisely an FP32
    result is defined at all."""
    from .plan import Plan
    rng = np.random.default_rng(seed)
    out = np.zeros((len(cases), trials + 1))
    for j, (kw, f0) in enumerate(cases):
        f0 = np.asarray(f0, dtype=np.float64)
        pl = Plan(**{**kw, "precision": 64}).set_field(f0).step(steps)
        ref = pl.reduce()[0]
        pl.close()
        for t in range(trials + 1):
            f = f0 if t == 0 else f0 + np.spacing(f0.astype(np.float32)).astype(np.float64) * rng.choice(
                [-1.0, 0.0, 1.0], size=f0.shape)
            pl = Plan(**{**kw, "precision": 32}).set_field(f).step(steps)
            out[j, t] = abs(pl.reduce()[0] - ref) / abs(ref)
            pl.close()
    return out


class BatchedSweep:
    """Instances grouped by grid size, one Batch per size.

    By default each size's Batch steps with its own (autotuned) launches on
    its own HIP stream, replaying 32-step graphs, so the sizes' launches
    overlap on the device.  grouped=True advances every size through one
    Group instead: per step one row and one column launch per register class
    over every size's tiles (variant-0 schedules); on the 256-instance sweep
    this measured slower, so it is not the default.  grouped="flow" steps the
    Group in its persistent form: one launch per register class per
    `flow_steps` steps over a work list of every size's tiles, each instance
    following its own chain of passes (pfc_group_set_flow)."""

    def __init__(self, instances, dt=0.5, precision=32, scheme="etd1", grouped=False, flow_steps=256):
        import torch
        from .plan import Batch
        self.instances = list(instances)
        self.groups = {}
        for j, inst in enumerate(self.instances):
            self.groups.setdefault(inst.N, []).append(j)
        # the seeded initial fields, drawn on host threads (NumPy's generators
        # release the GIL; each field has its own seed, so the values do not
        # depend on the thread count)
        from concurrent.futures import ThreadPoolExecutor
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
        with ThreadPoolExecutor(workers) as ex:
            fields = list(ex.map(initial_field, self.instances))
        self.batches, self.streams = {}, {}
        for n, idx in self.groups.items():
            first = self.instances[idx[0]]
            b = Batch(len(idx), **_model_kw(first, dt, precision, scheme))
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for slot, j in enumerate(idx):
                    inst = self.instances[j]
                    if (inst.Lx, inst.Ly) != (first.Lx, first.Ly):
                        raise ValueError("instances of one grid size must share the box")
                    b.set_slot(slot, inst.r)
                    b.set_field(slot, fields[j])
            self.batches[n], self.streams[n] = b, s
        torch.cuda.synchronize()
        self.group = None
        if grouped:
            from .plan import Group
            self.group = Group([self.batches[n] for n in sorted(self.batches, reverse=True)],
                               flow=flow_steps if grouped == "flow" else 0)

    def step(self, steps):
        import torch
        if self.group is not None:
            self.group.step(steps)
            return
        for n, b in self.batches.items():
            with torch.cuda.stream(self.streams[n]):
                b.step(steps)

    def results(self):
        import torch
        if self.group is not None:
            torch.cuda.synchronize()  # the group stepped on the current stream
        m = len(self.instances)
        fe, mean, bad = np.zeros(m), np.zeros(m), np.zeros(m, dtype=np.int32)
        for n, b in self.batches.items():
            with torch.cuda.stream(self.streams[n]):
                red = b.reduce()
                st = b.status()
            for slot, j in enumerate(self.groups[n]):
                fe[j], mean[j] = red[slot, 0], red[slot, 1]
                bad[j] = st[slot]
        return fe, mean, bad

    def close(self):
        if self.group is not None:
            self.group.close()
        for b in self.batches.values():
            b.close()


def run_batched(instances, steps, dt=0.5, precision=32, scheme="etd1", chunk=None, grouped=False, flow_steps=256):
    """All instances at once; the wall time covers setup (allocation, tables,
    upload, kernel tuning), stepping and the reductions, as run_plain's does."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sw = BatchedSweep(instances, dt=dt, precision=precision, scheme=scheme, grouped=grouped, flow_steps=flow_steps)
    done = 0
    chunk = chunk or steps
    while done < steps:
        k = min(chunk, steps - done)
        sw.step(k)
        done += k
    fe, mean, bad = sw.results()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sw.close()
    return SweepResult(fe, mean, wall, len(instances) / wall, bad)


def coexistence_by_r(instances, fe):
    """Liquid-crystal coexistence (common tangent of the analytic liquid and
    the tabulated relaxed crystals) at every r among `instances`, `fe[j]`
    being the free energy of `instances[j]`: {r: Tangent}."""
    from .coexistence import std_sweep_coexistence
    groups = {}
    for inst, f in zip(instances, fe):
        groups.setdefault(inst.r, []).append((inst.psibar, f))
    return {r: std_sweep_coexistence(r, [p for p, _ in g], [f for _, f in g]) for r, g in sorted(groups.items())}


def maxwell_lines_by_r(instances, fe, crystal_margin=1e-6, min_points=4):
    """Maxwell lines of the sweep with r as the temperature: at every r the
    lower envelope of the analytic liquid and the tabulated crystal branch
    (the relaxed states at least `crystal_margin` below the liquid; r values
    with fewer than `min_points` of them have the liquid alone), both over
    the sweep's psibar range.  Returns a
    MaxwellLines, whose write_csv gives the rows / metastable / windows tables."""
    from .coexistence import std_liquid
    from .phase_diagram.common_tangent import lower_envelope, maxwell_lines, tabulated_branch
    groups = {}
    for inst, f in zip(instances, fe):
        groups.setdefault(inst.r, []).append((inst.psibar, f))
    envs = {}
    for r, g in sorted(groups.items()):
        psi = np.array([p for p, _ in g], dtype=float)
        f = np.array([v for _, v in g], dtype=float)
        # the liquid over the sweep's own densities: beyond them no crystal
        # data exist, and a longer liquid would add domain-limited bridges
        liq = std_liquid(r, float(psi.min()), float(psi.max()))
        crystal = f < np.asarray(liq.f(psi)) - crystal_margin
        branches = [liq]
        if crystal.sum() >= min_points:
            branches.append(tabulated_branch("crystal", psi[crystal], f[crystal], min_points=min_points))
        envs[float(r)] = lower_envelope(branches, temp=float(r))
    return maxwell_lines(envs)
