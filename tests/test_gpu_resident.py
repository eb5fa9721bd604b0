"""The resident stepper (small ETD1 instances, every step of a call in one
workgroup's LDS) against the two-kernel step (needs a GPU).

It restates k_row_nonlin and k_col_etd's arithmetic; the compiler contracts
multiply-adds into FMAs per kernel, so against the generic kernels
(PFC_HIP_GENERIC=1, PFC_HIP_RESIDENT=0) the fields agree to round-off
(<= 1e-12 FP64, <= 2e-5 FP32 after 300 steps on O(1) fields).  Between plans
and batch slots, and between chunked and single-step calls, which run the
same kernel, they are identical bit for bit."""
import os

import numpy as np
import pytest

from pfc_hip import fields as F

pytestmark = pytest.mark.gpu


def _plan(resident, **kw):
    from pfc_hip.plan import Plan
    saved = {k: os.environ.get(k) for k in ("PFC_HIP_RESIDENT", "PFC_HIP_GENERIC")}
    try:
        if not resident:
            os.environ["PFC_HIP_RESIDENT"] = "0"
            os.environ["PFC_HIP_GENERIC"] = "1"
        return Plan(**kw)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


CASES = [
    dict(model="std", Nx=28, Ny=24, epsilon=-0.25, g=0.0, mean=-0.25, dt=0.5, precision=64),
    dict(model="log", Nx=28, Ny=24, epsilon=-0.9, g=-2.0, mean=0.63, dt=0.1, precision=64),
    dict(model="std", Nx=30, Ny=30, epsilon=-0.3, g=0.5, mean=-0.2, dt=0.3, precision=64),
    dict(model="std", Nx=16, Ny=10, epsilon=-0.3, g=0.0, mean=-0.2, dt=0.3, precision=64),
    dict(model="log", Nx=28, Ny=24, epsilon=-0.9, g=-2.0, mean=0.63, dt=0.1, precision=32),
    dict(model="std", Nx=40, Ny=36, epsilon=-0.25, g=0.0, mean=-0.25, dt=0.5, precision=32),
]


def _kw(c):
    a = 4 * np.pi / 3 ** 0.5
    return dict(model=c["model"], Nx=c["Nx"], Ny=c["Ny"], Lx=2 * a, Ly=4 * np.pi, Gamma=1.0,
                beta=0.5 if c["model"] == "log" else 1.0, q0=1.0, epsilon=c["epsilon"], g=c["g"], dt=c["dt"],
                precision=c["precision"])


@pytest.mark.parametrize("c", CASES, ids=lambda c: f'{c["model"]}{c["Nx"]}x{c["Ny"]}fp{c["precision"]}')
def test_resident_equals_generic_kernels(hip_device, c):
    x0 = F.uniform_noise((c["Ny"], c["Nx"]), c["mean"], 0.03, seed=11)
    res = _plan(True, **_kw(c)).set_field(x0)
    gen = _plan(False, **_kw(c)).set_field(x0)
    assert res.schedule()["row"][0] == -1 and gen.schedule()["row"][0] == 0
    res.step(300)
    gen.step(300)
    a, b = res.get_field().cpu().numpy(), gen.get_field().cpu().numpy()
    fa, fb = res.reduce()[0], gen.reduce()[0]
    print(f"\n{c['model']} {c['Nx']}x{c['Ny']} fp{c['precision']}: max |d| {np.abs(a - b).max():.3e} "
          f"df {abs(fa - fb):.3e}", flush=True)
    tol = 1e-12 if c["precision"] == 64 else 2e-5
    assert np.abs(a - b).max() <= tol
    assert abs(fa - fb) <= (1e-13 if c["precision"] == 64 else 1e-6) * max(1.0, abs(fb))
    assert res.status() == gen.status() == 0


def test_resident_chunks_and_single_steps_agree(hip_device):
    c = CASES[1]
    x0 = F.uniform_noise((c["Ny"], c["Nx"]), c["mean"], 0.03, seed=12)
    one = _plan(True, **_kw(c)).set_field(x0).step(5000)       # two launches: 4096 + 904 steps
    many = _plan(True, **_kw(c)).set_field(x0)
    for _ in range(50):
        many.step(100)
    assert np.array_equal(one.get_field().cpu().numpy(), many.get_field().cpu().numpy())


def test_resident_batch_slots_equal_plans(hip_device):
    # slots with their own epsilon, box and dt, stepped in one launch, equal
    # separate plans; an inactive slot is untouched
    from pfc_hip.plan import Batch
    c = CASES[1]
    kw = _kw(c)
    setups = [(-0.9, 0.95, 0.1), (-0.7, 1.0, 0.05), (-0.9, 1.05, 0.1), (-1.1, 0.9, 0.02)]
    fields = [F.uniform_noise((c["Ny"], c["Nx"]), 0.63, 0.03, seed=20 + i) for i in range(len(setups) + 1)]
    b = Batch(len(setups) + 1, **kw)
    for i, (eps, q, dt) in enumerate(setups):
        b.set_slot(i, eps)
        b.set_slot_box(i, kw["Lx"] / q, kw["Ly"] / q, dt)
    for i, f in enumerate(fields):
        b.set_field(i, f)
    b.set_active(range(len(setups))).step(400)
    for i, (eps, q, dt) in enumerate(setups):
        p = _plan(True, **{**kw, "epsilon": eps, "Lx": kw["Lx"] / q, "Ly": kw["Ly"] / q, "dt": dt})
        p.set_field(fields[i]).step(400)
        assert np.array_equal(b.get_field(i).cpu().numpy(), p.get_field().cpu().numpy()), i
    idle = b.get_field(len(setups)).cpu().numpy()
    ref = _plan(True, **kw).set_field(fields[-1]).get_field().cpu().numpy()
    assert np.array_equal(idle, ref)
