"""The 1024^2 seeded-crystal ETDRK2 workload at its own length (needs a GPU).

stdPFC, 1024 x 1024, dx = pi/4, r = psibar = -0.3, g = 0, ETDRK2 dt = 0.5,
20 000 steps from psibar + N(0, 0.01^2) with 8 one-mode triangular seeds
(seed 2).  The reference: the FP64 oracle's trajectory in
tests/golden/config2_1024_etdrk2_oracle.json (tests/golden/make_config2_golden.py:
free energy and mean every 250 steps; the 12 strongest structure-factor bins,
min, max and a 16 x 16 sample of the field at steps 100, 2000, 5000, 10 000
and 20 000).

* FP64: relative free energy per site <= 1e-9 at every record, the mean to
  1e-12, the same structure-factor peak bins at every checkpoint, and the
  field sample to 1e-10 after 100 steps;
* FP32 bulk (FP64 tables, combine and reductions): relative free energy
  <= 1e-6 at every record, the same peak bins, and the mean held by the k = 0
  pin to FP32 rounding.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD_PATH = os.path.join(HERE, "golden", "config2_1024_etdrk2_oracle.json")


def _gold():
    g = json.load(open(GOLD_PATH))
    if not g["records"] or g["records"][-1]["step"] < 20000:
        pytest.skip("oracle trajectory incomplete")
    return g


def _run(precision):
    from pfc_hip.plan import Plan
    n = 1024
    L = n * math.pi / 4
    geom = O.build_geometry(n, n, L, L)
    psi0 = F.seeded_crystals(geom, -0.3, -0.3, n_seeds=8, radius_pts=20, noise_std=0.01, seed=2)
    g = _gold()
    recs = {r["step"]: r for r in g["records"]}
    cps = {int(k): v for k, v in g["checkpoints"].items()}
    pl = Plan(Nx=n, Ny=n, Lx=L, Ly=L, epsilon=-0.3, g=0.0, dt=0.5, scheme="etdrk2", precision=precision,
              pin_mean=True)
    pl.set_field(psi0)
    done, fe_err, mean_err, out = 0, 0.0, 0.0, {}
    for target in sorted(set(recs) | set(cps)):
        pl.step(target - done)
        done = target
        fe, mean, lo, hi = pl.reduce()
        ref = recs.get(target, cps.get(target))
        fe_err = max(fe_err, abs(fe - ref["free_energy"]) / abs(ref["free_energy"]))
        if target in recs:
            mean_err = max(mean_err, abs(mean - recs[target]["mean"]))
        if target in cps:
            f = pl.get_field().cpu().numpy()
            out[target] = (O.structure_factor_peaks(f, 12), f[::64, ::64], lo, hi)
    print(f"config2 FP{precision}: max rel FE {fe_err:.3e}, max |d mean| {mean_err:.3e} over {done} steps",
          flush=True)
    assert pl.status() == 0
    return g, cps, fe_err, mean_err, out


def test_fp64_20000_steps_against_the_oracle(hip_device):
    g, cps, fe_err, mean_err, out = _run(64)
    assert fe_err <= 1e-9
    assert mean_err <= 1e-12
    for step, (peaks, sample, lo, hi) in out.items():
        assert [tuple(p) for p in cps[step]["peaks12"]] == peaks, step
        d = float(np.abs(sample - np.array(cps[step]["sample16"])).max())
        print(f"config2 FP64 step {step}: sample max|dpsi| {d:.2e}", flush=True)
        if step == 100:
            assert d <= 1e-10


def test_fp32_20000_steps_against_the_oracle(hip_device):
    g, cps, fe_err, mean_err, out = _run(32)
    assert fe_err <= 1e-6
    assert mean_err <= 1e-7
    for step, (peaks, sample, lo, hi) in out.items():
        assert [tuple(p) for p in cps[step]["peaks12"]] == peaks, step
