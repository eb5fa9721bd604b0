#!/usr/bin/env python3
"""Benchmark: the stdPFC ETDRK2 step on a 1024^2 grid (one MI355X).

Workload (the standard-PFC single-instance configuration): 1024 x 1024,
dx = dy = pi/4, r = epsilon = -0.3, psibar = -0.3, g = 0, ETDRK2 with dt = 0.5;
initial field psibar + N(0, 0.01^2) holding 8 one-mode crystal seeds of radius
20 grid points (seeded, synthetic).  The field and tables are resident in HBM
before the timed region.  A "step" is one ETDRK2 time step (two nonlinear
evaluations, four kernels).

Prints ONE JSON line.  `value` is Mpts*steps/s of the plain FP64 step;
`fp32` carries the FP32-bulk / FP64-reduction variant of the same workload.
`roofline` is measured for the dominant kernel (largest total time) with a HIP
event recorded after every launch on the stream the kernels run on, launches
back to back; `cpu_baseline` times the FP64 NumPy/SciPy CPU restatement of the
same step on this host's cores over a bounded sample.  `sweep`,
`phase_diagram` and `logpfc` carry the 256-instance sweep, the batched q0*
scans and the logPFC 2048^2 stabilised workload (`--sweep`, `--phase-diagram`,
`--logpfc` run one of them alone).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pfc-models_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
N = 1024
R = PSIBAR = -0.3
DT = 0.5


def _workload():
    import numpy as np
    from pfc_hip import fields as F
    from pfc_hip.discretization import Build2DGeometry
    L = N * math.pi / 4
    geom = Build2DGeometry(N, N, L, L)
    psi0 = F.seeded_crystals(geom, R, PSIBAR, n_seeds=8, radius_pts=20, noise_std=0.01, seed=2)
    return geom, np.ascontiguousarray(psi0)


def _algorithmic_bytes(kind, precision):
    """Bytes a kernel launch must move at minimum: half spectra of Nc complex
    read and written once each (row pass: S~ in, T out; column pass ETD/stage 1:
    S~ and T in, S~ and the k-space carry out; stage 2: T and carry in, S~ out)."""
    c = 16 if precision == 64 else 8
    nc = N * (N // 2 + 1)
    return {"row": 2, "col": 4, "col2": 3}[kind] * nc * c


def run_gpu(precision, steps, warmup):
    import torch
    from pfc_hip.plan import Plan
    geom, psi0 = _workload()
    pl = Plan(Nx=N, Ny=N, Lx=geom.Lx, Ly=geom.Ly, Gamma=1.0, beta=1.0, q0=1.0, epsilon=R, g=0.0, dt=DT,
              precision=precision, scheme="etdrk2", pin_mean=True)
    pl.set_field(psi0)
    pl.step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pl.step(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fe, mean, _, _ = pl.reduce()
    graphs = pl.graph_stats()
    # the same steps launched directly (calls of < 32 steps replay no graph),
    # for the step-loop graph's gain; the state goes on from the timed run
    t0 = time.perf_counter()
    left = steps
    while left > 0:
        pl.step(min(left, 31))
        left -= 31
    torch.cuda.synchronize()
    eager_wall = time.perf_counter() - t0
    # per-kernel device time: each kernel kind of the schedule in use, launched
    # back to back on scratch copies of the live state, one HIP event pair per
    # kind on the plan's stream; weighted by launches per ETDRK2 step (two row
    # passes, one of each column pass) to pick the dominant kernel
    nprof = min(steps, 500)
    prof = pl.step_profile(nprof)
    per_step = {"row": 2, "col": 1, "col2": 1}
    kinds = {k: (ms / calls, per_step[k]) for k, (ms, calls) in prof.items() if calls}
    kernel_ms_per_step = sum(avg * n for avg, n in kinds.values())
    dom = max(kinds, key=lambda k: kinds[k][0] * kinds[k][1])
    avg_ms = kinds[dom][0]
    ab = _algorithmic_bytes(dom, precision)
    achieved = ab / (avg_ms * 1e-3) / 1e9
    sched = pl.schedule()
    rk = "k_row_fast" if sched["row"][0] else "k_row_nonlin"
    ck = "k_col_fast" if sched["col"][0] else "k_col_etd"
    names = {"row": rk, "col": f"{ck}<COL_RK2_STAGE1>", "col2": f"{ck}<COL_RK2_STAGE2>"}
    step_bytes = 2 * 4 * N * (N // 2 + 1) * (16 if precision == 64 else 8)
    traffic = _traffic(names[dom], precision)
    return {
        "value": N * N * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "eager_ms_per_step": eager_wall / steps * 1e3,
        "graph_captures_replays": list(graphs),
        "free_energy": fe,
        "mean": mean,
        "kernels_us": {k: round(v[0] * 1e3, 3) for k, v in kinds.items()},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": names[dom],
            "algorithmic_bytes_per_launch": ab, "avg_launch_us": avg_ms * 1e3,
            "kernels_ms_per_step": kernel_ms_per_step,
        },
        "step_roofline": {
            "bytes_per_step": step_bytes,
            "achieved": step_bytes / (wall / steps) / 1e9,
            "frac": step_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
        },
        "device_bytes": pl.device_bytes(),
        "schedule": sched,
        # the same kernels under PFC_HIP_SCHEDULE64 / PFC_HIP_SCHEDULE32
        "schedule_pin": f"{sched['row'][1]}:{sched['row'][3]}:{sched['col'][1]}:{sched['col'][3]}",
        "stores": pl.stores(),
    }


def _traffic(kernel, precision):
    """HBM bytes per launch of `kernel` from the committed counter record
    (rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled by the same
    counter on a plain copy of known size; every schedule variant measured
    agreed within 1 %), when it covers this kernel and precision; else None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{N}sq_fp{precision}.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("hbm_bytes_per_launch", {}).get(kernel)


def run_copy_calibration(reps=50):
    """A plain device copy of one FP64 row-spectrum array (the row pass's read
    footprint), for calibrating rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    nbytes = N * (N // 2 + 1) * 16
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": nbytes, "bytes_written": nbytes, "launches": reps}}),
          flush=True)


def run_serve(n=2000, load=0.8, sample_per_side=2, capacity_sample=600):
    """The served workload: `n` seeded requests (70 % stdPFC, 30 % logPFC,
    sides 128..2048, 500..20000 ETD1 steps, FP32 bulk), HBM budget 200 GB.
    First the stream's first `capacity_sample` requests are queued at once
    (capacity, solves/s); then all `n` arrive as a Poisson stream at `load` x
    that capacity and the time-to-result (arrival to result on the host) is
    reported.  Accuracy: a sample of served results against the plain FP64
    one-at-a-time solve."""
    import numpy as np
    import torch
    from pfc_hip import serve as SV
    from pfc_hip.plan import Plan

    def note(tag):
        return lambda d, tot, t: print(f"serve[{tag}]: {d}/{tot} done at {t:.0f} s", file=sys.stderr, flush=True)

    srv = SV.Server(hbm_budget=200e9)
    ncap = min(n, capacity_sample)
    cap = srv.run(SV.poisson_stream(n=n, rate=1.0)[:ncap], realtime=False, progress=note("capacity"))
    srv.close()
    rate = load * cap.solves_per_s
    srv = SV.Server(hbm_budget=200e9)
    live = srv.run(SV.poisson_stream(n=n, rate=rate), realtime=True, progress=note("poisson"))
    srv.close()
    assert live.rejected == 0 and len(live.results) == n
    # accuracy: two live results per grid side against the plain FP64
    # one-at-a-time solve
    reqs = {r.id: r for r in SV.poisson_stream(n=n, rate=rate)}
    rng = np.random.default_rng(11)
    ids = []
    for side in sorted({q.N for q in reqs.values()}):
        pool = [r.id for r in live.results if reqs[r.id].N == side]
        ids += list(rng.choice(pool, size=min(sample_per_side, len(pool)), replace=False))
    rel = []
    for i in ids:
        q = reqs[int(i)]
        pl = Plan(**SV.request_params(q, precision=64))
        pl.set_field(SV.request_field(q))
        pl.step(q.steps)
        fe = pl.reduce()[0]
        pl.close()
        rel.append(abs(live.results[int(i)].fe - fe) / abs(fe))
        print(f"serve[check]: request {int(i)} N={q.N} {q.model} steps={q.steps} rel FE {rel[-1]:.2e}",
              file=sys.stderr, flush=True)
    # the plain FP64 one-at-a-time rate on the whole mix: each (side, model)'s
    # FP64 step time, measured on its own plan, times every request's steps
    t_step = {}
    for key in sorted({(q.N, q.model) for q in reqs.values()}):
        q = next(r for r in reqs.values() if (r.N, r.model) == key)
        pl = Plan(**SV.request_params(q, precision=64))
        pl.set_field(SV.request_field(q))
        pl.step(20)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pl.step(200)
        pl.reduce()
        t_step[key] = (time.perf_counter() - t0) / 200
        pl.close()
    plain_total = sum(q.steps * t_step[(q.N, q.model)] for q in reqs.values())
    rel = np.array(rel)
    return {
        "metric": f"served time-to-result ({n} requests, 70/30 std/log, sides 128..2048, 500..20000 ETD1 steps, FP32)",
        "capacity_solves_per_s": round(cap.solves_per_s, 3),
        "capacity_basis": f"the stream's first {ncap} requests queued at once",
        "capacity_wall_s": round(cap.wall_s, 2),
        "poisson_rate_per_s": round(rate, 3),
        "solves_per_s": round(live.solves_per_s, 3),
        "p50_s": round(live.p50_s, 3),
        "p99_s": round(live.p99_s, 3),
        "wall_s": round(live.wall_s, 2),
        "occupied_fraction": round(live.occupied_fraction, 3),
        "slabs": {"created": live.slabs_created, "released": live.slabs_released, "drained": live.slabs_drained,
                  "peak_bytes": int(live.peak_bytes)},
        "shapes": len(live.pools),
        "nonfinite": int(sum(r.nonfinite for r in live.results)),
        "check_sample": f"{len(ids)} served results ({sample_per_side} per grid side) against plain FP64 one-at-a-time",
        "rel_fe_err": {"median": float(np.median(rel)), "max": float(rel.max())},
        "plain_fp64_solves_per_s": round(n / plain_total, 3),
        "plain_fp64_basis": "per-step FP64 time of each (side, model) on its own plan x every request's steps",
        "capacity_vs_plain_fp64": round(cap.solves_per_s / (n / plain_total), 2),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_cpu(max_seconds=20.0):
    """FP64 CPU restatement of the same step (scipy.fft), on all host cores
    (up to 16, the value) and on one core (`single_core`, a shorter sample)."""
    sys.path.insert(0, ROOT)
    from oracle import pfc_oracle as O
    geom, psi0 = _workload()
    og = O.build_geometry(N, N, geom.Lx, geom.Ly)

    def timed(workers, seconds):
        sim = O.Simulation(O.Params(model="std", epsilon=R, g=0.0, dt=DT, scheme="etdrk2"), og,
                           fft="scipy", workers=workers)
        sim.set_field(psi0)
        sim.step(2)  # warm plans and caches
        done, t0 = 0, time.perf_counter()
        while done < 200:
            sim.step(5)
            done += 5
            if time.perf_counter() - t0 > seconds:
                break
        return done, time.perf_counter() - t0

    workers = min(len(os.sched_getaffinity(0)), 16)
    done, wall = timed(workers, max_seconds)
    done1, wall1 = timed(1, max_seconds / 2)
    return {"value": N * N * done / wall / 1e6, "unit": "Mpts*steps/s", "cores": workers, "kind": "port",
            "sample": f"{done} ETDRK2 FP64 steps of the 1024^2 seeded-crystal workload "
                      f"(NumPy/SciPy restatement of the reference CPU step, scipy.fft workers={workers})",
            "single_core": {"value": N * N * done1 / wall1 / 1e6, "cores": 1, "sample": f"{done1} steps, workers=1"},
            "cpu_model": _cpu_model()}


def run_sweep(plain_fraction=8, steps=5000):
    """The 256-instance (r, psibar) stdPFC sweep, 5000 ETD1 steps each: batched
    FP32 (and FP64) against the plain FP64 one-job-at-a-time step, the latter
    timed on every `plain_fraction`-th instance and reported per solve."""
    import numpy as np
    from pfc_hip import sweep as SW
    inst = SW.sweep_instances()
    # untimed 10-step runs of each form first: the process loads the sweep
    # sizes' kernel code and the autotuner caches their schedules; the timed
    # runs still include allocation, tables, fields, graph capture and results
    for prec, grouped in ((32, False), (64, False), (32, True), (32, "flow")):
        SW.run_batched(inst, 10, precision=prec, grouped=grouped)
    b32 = SW.run_batched(inst, steps, precision=32)
    b64 = SW.run_batched(inst, steps, precision=64)
    # the same sweep through one Group (a row and a column launch per
    # register class per step over every size's tiles)
    g32 = SW.run_batched(inst, steps, precision=32, grouped=True)
    # and through the Group's persistent form (one launch per register class
    # per 256 steps over the tiles of every size, per-instance dependencies)
    f32 = SW.run_batched(inst, steps, precision=32, grouped="flow")
    sub = inst[::plain_fraction]
    plain = SW.run_plain(sub, steps, checkpoint=min(500, steps))
    idx = [i.index for i in sub]
    rel32 = np.abs(b32.fe[idx] - plain.fe) / np.abs(plain.fe)
    rel64 = np.abs(b64.fe[idx] - plain.fe) / np.abs(plain.fe)
    # relaxations still moving by more than 1e-6 relative over the last 500
    # FP64 steps are reported apart: FP32 rounding shifts their timing
    relaxed = np.abs(plain.fe - plain.fe_prev) <= 1e-6 * np.abs(plain.fe)
    # coexistence densities at every r: FP32 batched against FP64 batched
    # (which matches the plain FP64 solve to rounding, rel64 above)
    # the floors the still-relaxing instances' FP32 errors are read against:
    # the FP64 solve's change under a one-ulp and an FP32-rounded start, and
    # the FP32 path's own spread over starts moved by up to one FP32 ulp
    moving = [i for i, ok in zip(sub, relaxed) if not ok]
    floors = None
    if moving:
        cases = [(SW._model_kw(i, 0.5, 64, "etd1"), SW.initial_field(i)) for i in moving]
        _, ic, ulp = SW.rounding_floor(cases, steps)
        spread = SW.fp32_spread(cases, steps)
        floors = {"fp64_one_ulp_start_max": float(ulp.max()), "fp64_fp32_rounded_start_max": float(ic.max()),
                  "fp32_spread_over_perturbed_starts": [[float(v.min()), float(v.max())] for v in spread]}
    c32, c64 = SW.coexistence_by_r(inst, b32.fe), SW.coexistence_by_r(inst, b64.fe)
    from pfc_hip.coexistence import liquid_crystal_pair_ok
    tangents = [r for r in c64 if c64[r].status == "ok" and c32[r].status == "ok"]
    both = [r for r in tangents if liquid_crystal_pair_ok(c64[r])]
    spurious = [r for r in tangents if r not in both]
    dpsi = [max(abs(c32[r].phi_a - c64[r].phi_a), abs(c32[r].phi_b - c64[r].phi_b)) for r in both]
    return {
        "metric": f"sweep solves/s (256 stdPFC instances, 64^2..512^2, ETD1 dt=0.5, {steps} steps)",
        "batched_fp32_solves_per_s": round(b32.solves_per_s, 2),
        "batched_fp64_solves_per_s": round(b64.solves_per_s, 2),
        "plain_fp64_solves_per_s": round(plain.solves_per_s, 2),
        "grouped_fp32_solves_per_s": round(g32.solves_per_s, 2),
        "persistent_fp32_solves_per_s": round(f32.solves_per_s, 2),
        "persistent_nonfinite_or_aborted": int(np.count_nonzero(f32.nonfinite)),
        "speedup_fp32_vs_plain_fp64": round(b32.solves_per_s / plain.solves_per_s, 2),
        "plain_sample": f"every {plain_fraction}th instance ({len(sub)} solves)",
        "timing": "whole job (allocation, tables, fields, graph capture, stepping, results) after an untimed "
                  "10-step run of the same sweep (kernel code loaded, schedules autotuned and cached)",
        "fp32_rel_fe_err": {"median": float(np.median(rel32)), "max": float(rel32.max())},
        "fp32_rel_fe_err_relaxed": {"n": int(relaxed.sum()),
                                    "median": float(np.median(rel32[relaxed])) if relaxed.any() else None,
                                    "max": float(rel32[relaxed].max()) if relaxed.any() else None},
        "still_relaxing": {"n": int((~relaxed).sum()), "max_rel_fe_err": float(rel32[~relaxed].max())
                           if (~relaxed).any() else None,
                           "rel_fe_err": [float(v) for v in rel32[~relaxed]], "floors": floors},
        "coexistence": {"r_values_with_tangent": len(tangents), "r_values_kept": len(both), "r_values": len(c64),
                        "spurious": {f"{r:.3f}": [round(c64[r].phi_a, 6), round(c64[r].phi_b, 6)] for r in spurious},
                        "spurious_rule": "phi_liquid >= phi_crystal - 1e-3 (reversed or degenerate: unconverged "
                                         "crystal free energies)",
                        "max_abs_density_err_fp32_vs_fp64": max(dpsi) if dpsi else None,
                        "fp64": {f"{r:.3f}": [round(c64[r].phi_a, 6), round(c64[r].phi_b, 6)] for r in both}},
        "fp64_batched_rel_fe_err_max": float(rel64.max()),
        "nonfinite": int((b32.nonfinite != 0).sum()),
    }


def run_logpfc(steps=10000, n=2048):
    """The logPFC 2048^2 workload: builder defaults (Gamma = beta = q0 = eps =
    g = 1, dt = 0.01, minPhi = 1e-6), dx = pi/4, psibar = -0.2 + U(-0.02,
    0.02) (seed 3) clipped to psi > -1 + 1e-6, ETD1 with the stabilised split
    (range check and M refit on the device before every step), `steps` steps
    in FP32 bulk and in FP64.  Mpts*steps/s over the steps after an untimed
    20-step start (tuning, first fit), and the final FP64 free energy against
    the uniform liquid's closed form at the same mean (the state this
    linearly stable phi = 0.8 relaxes to)."""
    import numpy as np
    import torch
    from pfc_hip import fields as F
    from pfc_hip.models.logpfc_hip import phase_diagram as LAD
    from pfc_hip.models.logpfc_hip import simulation as S
    L = n * math.pi / 4
    phi0 = 1.0 + np.clip(F.uniform_noise((n, n), -0.2, 0.02, seed=3), -1 + 1e-6, None)
    out = {"workload": f"logPFC {n}^2, dx=pi/4, builder defaults (dt=0.01), phi=0.8+U(-0.02,0.02), ETD1, "
                       f"stabilised split managed on the device, {steps} steps"}
    fes = {}
    for prec in (32, 64):
        geometry = S.Build2DGeometry(Nx=n, Ny=n, Lx=L, Ly=L)
        sim = S.BuildHipGPU_Simulation(S.BuildLogPFCModelParameters(meanDensity=float(phi0.mean())), geometry,
                                       precision=prec, stabilise=True)
        sim.State.phi[...] = phi0
        fe0 = float(sim.CalculateFreeEnergy())
        sim.run(20)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run(steps - 20)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        fe, mean, lo, hi = sim._plan.reduce()
        liq = LAD.liquid_free_energy(mean, {"beta": 1.0, "g": 1.0, "q0": 1.0, "v0": 1.0}, 1.0)
        fes[prec] = fe
        out[f"fp{prec}"] = {"value": n * n * (steps - 20) / wall / 1e6, "unit": "Mpts*steps/s",
                            "ms_per_step": 1e3 * wall / (steps - 20), "free_energy_start": fe0, "free_energy": fe,
                            "rel_above_liquid": (fe - liq) / abs(liq), "phi_min": lo, "phi_max": hi,
                            "M": float(sim.State.M), "status": int(sim.status()), "stores": sim._plan.stores()}
        sim.close()
    out["fp32_rel_fe_vs_fp64"] = abs(fes[32] - fes[64]) / abs(fes[64])
    return out


def run_sweep_forms(steps=2000):
    """The sweep's three forms on subsets of the 256 instances (FP32 ETD1,
    whole job as in the sweep leg, us per sweep step): per-size Batches on
    their own streams (autotuned, and with the autotuner off, i.e. the
    variant-0 schedules the group forms run), the per-pass Group, and the
    Group's persistent form (one launch per register class per `steps`
    steps).  It separates what the persistent form's rate owes to the
    launch structure from what it owes to the per-size schedule choice."""
    from pfc_hip import sweep as SW
    inst = SW.sweep_instances()
    subsets = {"all": inst, "class0": [i for i in inst if i.N in (64, 128, 256, 512)],
               "class1": [i for i in inst if i.N in (192, 384)], "512": [i for i in inst if i.N == 512],
               "256": [i for i in inst if i.N == 256]}
    forms = [("per_size_tuned", False, None), ("per_size_variant0", False, "0"), ("per_pass_group", True, None),
             ("persistent", "flow", None)]
    out = {"unit": "us per sweep step", "steps": steps}
    for name, sub in subsets.items():
        row = {"instances": len(sub)}
        for form, grouped, tune in forms:
            old = os.environ.get("PFC_HIP_TUNE")
            if tune is not None:
                os.environ["PFC_HIP_TUNE"] = tune
            try:
                SW.run_batched(sub, 10, precision=32, grouped=grouped, flow_steps=steps)
                r = SW.run_batched(sub, steps, precision=32, grouped=grouped, flow_steps=steps)
            finally:
                if tune is not None:
                    if old is None:
                        os.environ.pop("PFC_HIP_TUNE", None)
                    else:
                        os.environ["PFC_HIP_TUNE"] = old
            row[form] = round(r.wall_s / steps * 1e6, 1)
        out[name] = row
        print(f"sweep forms {name}: {row}", file=sys.stderr, flush=True)
    return out


def run_phase_diagram(n_dens=12, serial_every=1):
    """q0* over a density list with the phase-diagram driver: logPFC triangle,
    beta 0.5, g -2, eps -0.9 on the 28 x 24 cell with the driver tests'
    relaxation and q00-scan settings (3 passes x 21 q00 per density).  All
    densities scanned batched (pass p of every density in one batch) against
    the serial driver, one density and one q00 at a time, on every
    `serial_every`-th density (all by default); results must be identical.
    Each pass lasts as long as its slowest relaxation (edge q00 points whose
    seeded instability grows for up to ~10^6 steps before a poor-fit stop),
    which bounds the batched rate."""
    import numpy as np
    from pfc_hip.models.logpfc_hip import phase_diagram as LOG
    from pfc_hip.phase_diagram import batched as B
    from pfc_hip.phase_diagram import q0star as Q
    relax = Q.RelaxSettings(dt=0.1, totalT=1e6, stepT=100.0, divs=10, maxDepth=4, requiredFitness=0.99,
                            minimumA=1e-6)
    qscan = Q.QScanSettings(passes=3, Deltaq=0.02, q00_Steps=21, Deltaq_Refinement=0.05)
    m = Q.MinimizeFreeEnergy(LOG, dict(beta=0.5, g=-2.0), "triangle", dict(uNx=14, uNy=24, Mx=2, My=1),
                             dict(mode="L2H", A1=1e-5, A2=1e-6, A3=1e-7), relax, qscan)
    rhos = np.round(np.linspace(0.54, 0.76, n_dens), 6)
    t0 = time.perf_counter()
    rb = B.scan_density_batched(m, -0.9, rhos)
    tb = time.perf_counter() - t0
    sub = rhos[::serial_every]
    t0 = time.perf_counter()
    rs = Q.scan_density(m, -0.9, sub)
    ts = time.perf_counter() - t0
    same = True
    for k in range(len(sub)):
        j = k * serial_every
        same &= str(rb["status"][j]) == str(rs["status"][k])
        same &= bool(np.isnan(rs["q0star"][k])
                     or (rb["q0star"][j] == rs["q0star"][k] and rb["fmin"][j] == rs["fmin"][k]))
    n_eval = int(rb["qscan"]["q00"].size)
    return {
        "metric": "q0* density scans per second (logPFC triangle, 28 x 24 cell, eps -0.9, 3 x 21 q00 relaxations)",
        "batched_scans_per_s": round(len(rhos) / tb, 3),
        "serial_scans_per_s": round(len(sub) / ts, 3),
        "speedup_batched_vs_serial": round((len(rhos) / tb) / (len(sub) / ts), 2),
        "densities": [float(x) for x in rhos],
        "batched_wall_s": round(tb, 2), "serial_wall_s": round(ts, 2), "serial_sample": [float(x) for x in sub],
        "evaluations": n_eval, "time_steps": int(rb["qscan"]["steps"].sum()),
        "statuses": [str(x) for x in rb["status"]],
        "q0star": [None if np.isnan(x) else round(float(x), 10) for x in rb["q0star"]],
        "batched_equals_serial": bool(same),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    # >= 256 warm-up steps: the step loop's graph is captured by the first
    # call of 256 steps or more, so the capture falls in the warm-up and the
    # timed steps replay it
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 leg")
    ap.add_argument("--serve", action="store_true",
                    help="only run the serving leg (minutes) and print its JSON line")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-phase-diagram", action="store_true", help="skip the phase-diagram leg")
    ap.add_argument("--phase-diagram", action="store_true",
                    help="only run the phase-diagram leg and print its JSON line")
    ap.add_argument("--sweep", action="store_true",
                    help="only run the 256-instance sweep leg and print its JSON line")
    ap.add_argument("--sweep-steps", type=int, default=5000, help="steps per sweep instance (short profiling runs)")
    ap.add_argument("--sweep-forms", action="store_true",
                    help="only time the sweep's per-size, per-pass group and persistent forms on subsets "
                         "(2000 steps) and print their JSON line")
    ap.add_argument("--no-logpfc", action="store_true", help="skip the logPFC 2048^2 leg")
    ap.add_argument("--logpfc", action="store_true",
                    help="only run the logPFC 2048^2 stabilised leg and print its JSON line")
    ap.add_argument("--logpfc-steps", type=int, default=10000, help="steps of the logPFC leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (counter calibration) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        run_copy_calibration()
        return
    if a.serve:
        print(json.dumps({"serve": run_serve(n=a.serve_requests)}), flush=True)
        return
    if a.phase_diagram:
        print(json.dumps({"phase_diagram": run_phase_diagram()}), flush=True)
        return
    if a.sweep:
        print(json.dumps({"sweep": run_sweep(steps=a.sweep_steps)}), flush=True)
        return
    if a.sweep_forms:
        print(json.dumps({"sweep_forms": run_sweep_forms()}), flush=True)
        return
    if a.logpfc:
        print(json.dumps({"logpfc": run_logpfc(steps=a.logpfc_steps)}), flush=True)
        return
    if a.gpus != 1:
        print("bench.py: the benchmark runs on one GPU; --gpus above 1 is not supported "
              "(independent instances would shard one scheduler per GPU)", file=sys.stderr)
        sys.exit(2)
    f64 = run_gpu(64, a.steps, a.warmup)
    f32 = None if a.no_fp32 else run_gpu(32, a.steps, a.warmup)
    sweep = None if a.no_sweep else run_sweep()
    pd = None if a.no_phase_diagram else run_phase_diagram()
    logpfc = None if a.no_logpfc else run_logpfc(steps=a.logpfc_steps)
    serve = None if a.no_serve else run_serve(n=a.serve_requests)
    cpu = None if a.no_cpu else run_cpu(a.cpu_seconds)
    line = {
        "metric": "Mpts*steps/s at 1024^2 (stdPFC ETDRK2 step, FP64; FP32 in 'fp32')",
        "value": round(f64["value"], 1),
        "unit": "Mpts*steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(f64["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 1024^2 liquid + 8 one-mode crystal seeds)",
        "config": {"workload": "stdPFC 1024^2, dx=pi/4, r=-0.3, psibar=-0.3, ETDRK2 dt=0.5, plain FP64 step",
                   "grid": [N, N], "scheme": "etdrk2", "parallelism": "single GPU"},
        "roofline": f64["roofline"],
        "step_roofline": f64["step_roofline"],
        "kernels_us": f64["kernels_us"],
        "stores": f64["stores"],
        "schedule_pin": f64["schedule_pin"],
        "step_graph": {"eager_ms_per_step": round(f64["eager_ms_per_step"], 5),
                       "captures_replays": f64["graph_captures_replays"]},
        "fp32": None if f32 is None else {
            "value": round(f32["value"], 1), "ms_per_step": round(f32["ms_per_step"], 5),
            "eager_ms_per_step": round(f32["eager_ms_per_step"], 5),
            "dtype": "f32 bulk, f64 tables, combine and reductions", "roofline": f32["roofline"],
            "kernels_us": f32["kernels_us"], "stores": f32["stores"], "schedule_pin": f32["schedule_pin"],
            "rel_fe_vs_fp64": abs(f32["free_energy"] - f64["free_energy"]) / abs(f64["free_energy"])},
        "cpu_baseline": cpu,
        "sweep": sweep,
        "phase_diagram": pd,
        "logpfc": logpfc,
        "serve": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
