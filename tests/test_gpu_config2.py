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
* FP32 bulk (FP64 tables, combine and reductions): the same peak bins at
  every checkpoint, the mean held by the k = 0 pin to 1e-7, and the free
  energy against the trajectory's own sensitivity.  The crystals' growth
  and their later rearrangements (grain-boundary events near steps 7000,
  7750, 13 250 and 18 750) amplify perturbations of FP32 size: the FP64 step
  started from the initial field kicked by one FP32 ulp per site (random
  sign, seed 7) departs from the oracle by up to ~4e-6 relative at those
  events and by <= 1e-7 elsewhere.  That kicked FP64 run is made in the test
  (its free energies are the floor), and FP32 is held to <= max(1e-6, 3 x
  the floor's maximum) over the whole trajectory.  Where the floor is
  <= 1e-7 ("quiet" records) FP32's own rounding, accumulated over the run,
  is what remains, and it depends on the FFT schedule's rounding order:
  6.7e-7, 1.5e-6 and 3.6e-6 at the worst quiet record for three schedule
  variants (bit-identical results otherwise; measured).  The test runs the
  variant-0 schedules (PFC_HIP_TUNE=0: the autotuner's pick changes between
  runs) and holds quiet records to 5e-6.
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


def _initial():
    n = 1024
    L = n * math.pi / 4
    geom = O.build_geometry(n, n, L, L)
    return F.seeded_crystals(geom, -0.3, -0.3, n_seeds=8, radius_pts=20, noise_std=0.01, seed=2)


def _run(precision, psi0=None):
    from pfc_hip.plan import Plan
    n = 1024
    L = n * math.pi / 4
    if psi0 is None:
        psi0 = _initial()
    g = _gold()
    recs = {r["step"]: r for r in g["records"]}
    cps = {int(k): v for k, v in g["checkpoints"].items()}
    pl = Plan(Nx=n, Ny=n, Lx=L, Ly=L, epsilon=-0.3, g=0.0, dt=0.5, scheme="etdrk2", precision=precision,
              pin_mean=True)
    pl.set_field(psi0)
    done, fe_err, mean_err, out, rel = 0, 0.0, 0.0, {}, {}
    for target in sorted(set(recs) | set(cps)):
        pl.step(target - done)
        done = target
        fe, mean, lo, hi = pl.reduce()
        ref = recs.get(target, cps.get(target))
        rel[target] = abs(fe - ref["free_energy"]) / abs(ref["free_energy"])
        fe_err = max(fe_err, rel[target])
        if target in recs:
            mean_err = max(mean_err, abs(mean - recs[target]["mean"]))
        if target in cps:
            f = pl.get_field().cpu().numpy()
            out[target] = (O.structure_factor_peaks(f, 12), f[::64, ::64], lo, hi)
    print(f"config2 FP{precision}: max rel FE {fe_err:.3e}, max |d mean| {mean_err:.3e} over {done} steps",
          flush=True)
    assert pl.status() == 0
    return g, cps, fe_err, mean_err, out, rel


def test_fp64_20000_steps_against_the_oracle(hip_device):
    g, cps, fe_err, mean_err, out, rel = _run(64)
    assert fe_err <= 1e-9
    assert mean_err <= 1e-12
    for step, (peaks, sample, lo, hi) in out.items():
        assert [tuple(p) for p in cps[step]["peaks12"]] == peaks, step
        d = float(np.abs(sample - np.array(cps[step]["sample16"])).max())
        print(f"config2 FP64 step {step}: sample max|dpsi| {d:.2e}", flush=True)
        assert d <= (1e-10 if step == 100 else 1e-9), step


def test_fp32_20000_steps_against_the_oracle(hip_device, monkeypatch):
    monkeypatch.setenv("PFC_HIP_TUNE", "0")  # a fixed schedule: the rounding order is part of the result
    g, cps, fe_err, mean_err, out, rel = _run(32)
    psi0 = _initial()
    ulp = np.spacing(np.abs(psi0).astype(np.float32)).astype(np.float64)
    kick = psi0 + ulp * np.random.default_rng(7).choice([-1.0, 1.0], size=psi0.shape)
    floor = _run(64, kick)[5]
    quiet = [s for s in rel if floor[s] <= 1e-7]
    worst_quiet = max(rel[s] for s in quiet)
    print(f"config2 FP32: one-ulp-kick FP64 floor max {max(floor.values()):.2e}; FP32 max {fe_err:.2e}; "
          f"{len(quiet)}/{len(rel)} quiet records, FP32 max there {worst_quiet:.2e}", flush=True)
    assert len(quiet) >= len(rel) // 2
    assert worst_quiet <= 5e-6
    assert fe_err <= max(1e-6, 3 * max(floor.values()))
    assert mean_err <= 1e-7
    for step, (peaks, sample, lo, hi) in out.items():
        assert [tuple(p) for p in cps[step]["peaks12"]] == peaks, step
