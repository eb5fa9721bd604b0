"""The resident stepper (small ETD1 instances, every step of a call in one
workgroup's LDS) against the two-kernel step (needs a GPU).

Its FFT form restates k_row_nonlin and k_col_etd's arithmetic; its dense form
(cells of at most 512 half-spectrum points, Nx, Ny <= 32) evaluates the same
linear maps as cosine/sine sums and per-column circulant matrices (the
column step in FP64).  Against the generic kernels (PFC_HIP_GENERIC=1,
PFC_HIP_RESIDENT=0), and between the two forms, FP64 fields agree to
round-off (<= 1e-12 after 300 steps on O(1) fields).  FP32 fields are held
to the FP64 field of the same run instead (<= 5e-5 each): the dense form's
FP32 rounding differs from the FFT kernels' (measured on logPFC 28 x 24
after 300 steps: 5.3e-6 from FP64 against the FFT kernels' 1.7e-5 and
2.9e-5 on two seeds), so the two FP32 forms differ by about the sum of
their errors.  Between plans and
batch slots, and between chunked and single-step calls, which run the same
kernel, they are identical bit for bit."""


def _close(a, b, c, x0, steps, **plans):
    """FP64: a and b within 1e-12; FP32: each within 5e-5 of the FP64 dense
    resident field from x0 after `steps` steps."""
    if c["precision"] == 64:
        assert np.abs(a - b).max() <= 1e-12
        return
    ref = _plan(True, **{**_kw(dict(c, precision=64)), **plans}).set_field(x0).step(steps).get_field().cpu().numpy()
    assert np.abs(a.astype(np.float64) - ref).max() <= 5e-5
    assert np.abs(b.astype(np.float64) - ref).max() <= 5e-5
import os

import numpy as np
import pytest

from pfc_hip import fields as F

pytestmark = pytest.mark.gpu


def _plan(resident, dense=True, batch=0, **kw):
    from pfc_hip.plan import Batch, Plan
    saved = {k: os.environ.get(k) for k in ("PFC_HIP_RESIDENT", "PFC_HIP_GENERIC", "PFC_HIP_RESIDENT_DENSE")}
    try:
        if not resident:
            os.environ["PFC_HIP_RESIDENT"] = "0"
            os.environ["PFC_HIP_GENERIC"] = "1"
        if not dense:
            os.environ["PFC_HIP_RESIDENT_DENSE"] = "0"
        return Batch(batch, pin_mean=True, **kw) if batch else Plan(**kw)
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
    # odd and small lengths (radix-3 and radix-5 transforms; a 4 x 2 cell)
    dict(model="std", Nx=16, Ny=9, epsilon=-0.3, g=0.5, mean=-0.2, dt=0.3, precision=64),
    dict(model="log", Nx=12, Ny=25, epsilon=-0.9, g=-2.0, mean=0.63, dt=0.1, precision=64),
    dict(model="std", Nx=4, Ny=2, epsilon=-0.3, g=0.0, mean=-0.2, dt=0.3, precision=64),
    # odd Nx/2 with Ny at the dense form's compiled row bounds (16 and 32
    # rows; 15 angles)
    dict(model="std", Nx=30, Ny=16, epsilon=-0.3, g=0.5, mean=-0.2, dt=0.3, precision=64),
    dict(model="log", Nx=26, Ny=32, epsilon=-0.9, g=-2.0, mean=0.63, dt=0.1, precision=64),
]


def _dense(c):
    # the dense form's limits: (Nx/2 + 1) Ny <= 512, Nx/2 - 1 <= 15, Ny <= 32
    return (c["Nx"] // 2 + 1) * c["Ny"] <= 512 and c["Nx"] // 2 - 1 <= 15 and c["Ny"] <= 32


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
    assert res.schedule()["row"][0] == (-2 if _dense(c) else -1) and gen.schedule()["row"][0] == 0
    res.step(300)
    gen.step(300)
    a, b = res.get_field().cpu().numpy(), gen.get_field().cpu().numpy()
    fa, fb = res.reduce()[0], gen.reduce()[0]
    print(f"\n{c['model']} {c['Nx']}x{c['Ny']} fp{c['precision']}: max |d| {np.abs(a - b).max():.3e} "
          f"df {abs(fa - fb):.3e}", flush=True)
    _close(a, b, c, x0, 300)
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


@pytest.mark.parametrize("c", [c for c in CASES if _dense(c)],
                         ids=lambda c: f'{c["model"]}{c["Nx"]}x{c["Ny"]}fp{c["precision"]}')
def test_dense_equals_fft_resident(hip_device, c):
    x0 = F.uniform_noise((c["Ny"], c["Nx"]), c["mean"], 0.03, seed=31)
    den = _plan(True, **_kw(c)).set_field(x0)
    fft = _plan(True, dense=False, **_kw(c)).set_field(x0)
    assert den.schedule()["row"][0] == -2 and fft.schedule()["row"][0] == -1
    den.step(300)
    fft.step(300)
    a, b = den.get_field().cpu().numpy(), fft.get_field().cpu().numpy()
    fa, fb = den.reduce()[0], fft.reduce()[0]
    print(f"\ndense vs fft {c['model']} {c['Nx']}x{c['Ny']} fp{c['precision']}: max |d| "
          f"{np.abs(a - b).max():.3e} df {abs(fa - fb):.3e}", flush=True)
    _close(a, b, c, x0, 300)
    assert abs(fa - fb) <= (1e-13 if c["precision"] == 64 else 1e-6) * max(1.0, abs(fb))
    assert den.status() == fft.status() == 0


@pytest.mark.parametrize("precision", [64, 32])
def test_dense_pinned_batch_equals_fft_resident(hip_device, precision):
    # batch slots with their own epsilon, box and dt and the pinned k = 0
    # mode (the mean held at the set field's), dense against the FFT form;
    # the mean stays put
    c = dict(CASES[1], precision=precision)
    kw = _kw(c)
    setups = [(-0.9, 0.95, 0.1), (-0.7, 1.0, 0.05), (-1.1, 0.9, 0.02)]
    fields = [F.uniform_noise((c["Ny"], c["Nx"]), 0.6 + 0.02 * i, 0.03, seed=40 + i) for i in range(len(setups))]
    out = []
    for dense, prec in ((True, precision), (False, precision), (True, 64)):
        b = _plan(True, dense=dense, batch=len(setups), **{**kw, "precision": prec})
        for i, (eps, q, dt) in enumerate(setups):
            b.set_slot(i, eps)
            b.set_slot_box(i, kw["Lx"] / q, kw["Ly"] / q, dt)
            b.set_field(i, fields[i])
        b.step(250)
        out.append([b.get_field(i).cpu().numpy().astype(np.float64) for i in range(len(setups))])
        assert not b.status().any()
    for i in range(len(setups)):
        if precision == 64:
            assert np.abs(out[0][i] - out[1][i]).max() <= 1e-12, i
        else:  # each FP32 form against the FP64 dense batch
            assert np.abs(out[0][i] - out[2][i]).max() <= 5e-5, i
            assert np.abs(out[1][i] - out[2][i]).max() <= 5e-5, i
        assert abs(out[0][i].mean() - fields[i].mean()) <= (1e-14 if precision == 64 else 2e-7), i
