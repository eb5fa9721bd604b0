"""Write the oracle trajectory of the 1024^2 seeded-crystal ETDRK2 workload.

Workload: stdPFC, 1024 x 1024, dx = dy = pi/4, r = psibar = -0.3, g = 0,
ETDRK2 dt = 0.5, 20 000 steps; initial field psibar + N(0, 0.01^2) with 8
one-mode triangular seeds of radius 20 grid points (pfc_hip.fields.seeded_crystals,
seed 2).  The oracle (NumPy/SciPy FP64 restatement of the reference's CPU step)
records, every `every` steps, the free energy per site and the mean; at the
listed checkpoints also the 12 strongest structure-factor bins, the field's
min and max and a 16 x 16 sample of the field (rows and columns 0, 64, ...).

The GPU tests compare against this file; the oracle takes ~0.3 s per step on
six cores here, so the trajectory is computed once and committed:

    python tests/golden/make_config2_golden.py --workers 6

Progress is written to the output file after every record, so an interrupted
run can be continued with --resume (the state is kept in a side .npy file).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "pfc-models_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from oracle import pfc_oracle as O  # noqa: E402
from pfc_hip import fields as F  # noqa: E402

OUT = os.path.join(HERE, "config2_1024_etdrk2_oracle.json")


def workload():
    n = 1024
    L = n * math.pi / 4
    geom = O.build_geometry(n, n, L, L)
    psi0 = F.seeded_crystals(geom, -0.3, -0.3, n_seeds=8, radius_pts=20, noise_std=0.01, seed=2)
    p = O.Params(model="std", epsilon=-0.3, g=0.0, dt=0.5, scheme="etdrk2")
    return geom, p, np.ascontiguousarray(psi0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--checkpoints", default="100,2000,5000,10000,20000")
    ap.add_argument("--workers", type=int, default=6)
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--resume", action="store_true")
    a = ap.parse_args()
    geom, p, psi0 = workload()
    sim = O.Simulation(p, geom, fft="scipy", workers=a.workers)
    side = a.out + ".state.npy"
    cps = sorted(int(c) for c in a.checkpoints.split(","))
    if a.resume and os.path.exists(a.out) and os.path.exists(side):
        rec = json.load(open(a.out))
        sim.set_field(np.load(side))
        done = rec["records"][-1]["step"]
    else:
        sim.set_field(psi0)
        rec = {
            "workload": "stdPFC 1024^2, dx=pi/4, r=psibar=-0.3, g=0, ETDRK2 dt=0.5, seeded_crystals(seed=2)",
            "oracle": "oracle/pfc_oracle.py Simulation (scipy.fft, FP64)",
            "initial": {"free_energy": sim.free_energy(), "mean": O.exact_mean(psi0)},
            "records": [], "checkpoints": {},
        }
        done = 0
    t0 = time.time()
    steps_per_cp = set(cps)
    while done < a.steps:
        nxt = min(a.steps, (done // a.every + 1) * a.every)
        nxt = min([nxt] + [c for c in cps if c > done])
        sim.step(nxt - done)
        done = nxt
        if done % a.every == 0:
            rec["records"].append({"step": done, "free_energy": sim.free_energy(), "mean": O.exact_mean(sim.field)})
        if done in steps_per_cp:
            f = sim.field
            rec["checkpoints"][str(done)] = {
                "free_energy": sim.free_energy(),
                "peaks12": O.structure_factor_peaks(f, 12),
                "min": float(f.min()), "max": float(f.max()),
                "sample16": f[::64, ::64].tolist(),
            }
        np.save(side, sim.field)
        with open(a.out + ".tmp", "w") as fh:
            json.dump(rec, fh, indent=1)
        os.replace(a.out + ".tmp", a.out)
        print(f"step {done}  fe {sim.free_energy():.15f}  {time.time() - t0:.0f} s", flush=True)
    os.remove(side)


if __name__ == "__main__":
    main()
