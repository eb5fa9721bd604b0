"""HIP path against the FP64 oracle on identical seeded inputs (needs a GPU).

Tolerances (stated per test):
  FFT round trips            rel L2 <= 1e-14 (FP64), 3e-6 (FP32)
  FP64 step, 100 steps        max|d psi| <= 1e-10
  recorded FEs (FP64)         rtol 1e-10                    (the reference's own tolerance)
  free energy / mean (FP64)   rel 1e-12 against the oracle on the same field
"""
import numpy as np
import pytest

from oracle import pfc_oracle as O
from pfc_hip import fields as F

pytestmark = pytest.mark.gpu


def _plan(geom, precision=64, **kw):
    from pfc_hip.plan import Plan
    return Plan(Nx=geom.Nx, Ny=geom.Ny, Lx=geom.Lx, Ly=geom.Ly, precision=precision, **kw)


@pytest.mark.parametrize("shape", [(24, 28), (64, 64), (132, 132), (48, 96), (1024, 1024), (40, 90)])
def test_rfft2_irfft2_match_numpy(hip_device, shape):
    import torch
    from pfc_hip.plan import rfft2, irfft2
    rng = np.random.default_rng(11)
    x = rng.normal(size=(3, *shape))
    X = rfft2(torch.as_tensor(x, device="cuda")).cpu().numpy()
    ref = np.fft.rfft2(x)
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < 1e-14
    # inverse of a non-Hermitian-consistent spectrum: numpy drops the imaginary
    # parts of the kx=0 / Nyquist columns after the column inverse; so must we
    Z = ref + 1e-3 * (rng.normal(size=ref.shape) + 1j * rng.normal(size=ref.shape))
    z = irfft2(torch.as_tensor(Z, device="cuda"), shape).cpu().numpy()
    zr = np.fft.irfft2(Z, s=shape)
    assert np.linalg.norm(z - zr) / np.linalg.norm(zr) < 1e-14


def test_rfft2_fp32(hip_device):
    import torch
    from pfc_hip.plan import rfft2, irfft2
    rng = np.random.default_rng(12)
    x = rng.normal(size=(2, 256, 384)).astype(np.float32)
    X = rfft2(torch.as_tensor(x, device="cuda")).cpu().numpy()
    ref = np.fft.rfft2(x.astype(np.float64))
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < 3e-6
    back = irfft2(torch.as_tensor(X, device="cuda"), (256, 384)).cpu().numpy()
    assert np.linalg.norm(back - x) / np.linalg.norm(x) < 3e-6


def _std128(scheme="etd1"):
    geom = O.build_geometry(128, 128, 128 * np.pi / 4, 128 * np.pi / 4)
    p = O.Params(model="std", Gamma=1.0, beta=1.0, q0=1.0, epsilon=-0.25, g=0.0, dt=0.5, scheme=scheme)
    psi0 = F.uniform_noise((128, 128), -0.25, 0.05, seed=1)
    return geom, p, psi0


@pytest.mark.parametrize("scheme", ["etd1", "etdrk2"])
def test_fp64_step_matches_oracle_100_steps(hip_device, scheme):
    """max|d psi| <= 1e-10 after 100 FP64 steps (128^2, r = psibar = -0.25, dt = 0.5)."""
    geom, p, psi0 = _std128(scheme)
    sim = O.Simulation(p, geom)
    sim.set_field(psi0)
    sim.step(100)
    pl = _plan(geom, epsilon=-0.25, g=0.0, dt=0.5, scheme=scheme)
    pl.set_field(psi0)
    pl.step(100)
    got = pl.get_field().cpu().numpy()
    err = np.abs(got - sim.field).max()
    assert err <= 1e-10, err
    fe, mean, mn, mx = pl.reduce()
    assert fe == pytest.approx(sim.free_energy(), rel=1e-12)
    assert mean == pytest.approx(sim.mean(), rel=1e-13, abs=1e-15)
    assert mn == pytest.approx(sim.field.min(), abs=1e-10)
    assert mx == pytest.approx(sim.field.max(), abs=1e-10)


def test_recorded_stdpfc_free_energy(hip_device):
    """models/stdpfc_cpu/test_regression.py:64-115 on the HIP FP64 path (132^2 = (2^2 3 11)^2)."""
    a, h = 4 * np.pi / 3**0.5, 4 * np.pi
    c12, c34, c14 = (3/4)**0.5, (3/4)**0.75, (3/4)**0.25
    geom = O.build_geometry(132, 132, 14 * a, 8 * h)
    pl = _plan(geom, beta=c12 * 0.5, epsilon=c12 * (-0.9 + 2*(-2.0) + 4), g=c34 * (-2.0 + 2.5), dt=0.1)
    pl.set_field(F.GeneratePerfectCrystalInitialWithCentralVacancy(geom, c14*(0.55-1), c14*(0.30-1), A0=c14*0.19))
    m0 = pl.mean()
    pl.step(200)
    fe, mean, _, _ = pl.reduce()
    assert abs(mean - m0) < 1e-12
    assert fe == pytest.approx(-0.11237092135766821, rel=1e-10)


def test_recorded_logpfc_free_energies(hip_device):
    """models/logpfc_cpu/test_regression.py:28-66 on the HIP FP64 path (28 x 24 = 2^2 7 x 2^3 3)."""
    Q, MEAN = 0.9648639799649456, 0.6326
    a, h = 4 * np.pi / (np.sqrt(3) * Q), 4 * np.pi / Q
    geom = O.build_geometry(28, 24, 2 * a, h)
    X, Y = geom.X, geom.Y
    phi = MEAN * (1 + 0.15 * (2 * np.cos(np.sqrt(3) / 2 * Q * X) * np.cos(Q * Y / 2) + np.cos(Q * Y)))
    phi += 1e-3 * np.random.default_rng(7).standard_normal(phi.shape)
    phi += MEAN - phi.mean()
    pl = _plan(geom, model="log", beta=0.5, epsilon=-0.9, g=-2.0, dt=0.1)
    pl.set_field(phi)
    ref = {10: -0.5162767800170736, 100: -0.554845641895177, 1000: -0.5548456430520675}
    done = 0
    for k in sorted(ref):
        pl.step(k - done)
        done = k
        assert pl.free_energy() == pytest.approx(ref[k], rel=1e-10), k
    assert pl.mean() == pytest.approx(MEAN, abs=1e-12)


def test_reductions_match_oracle_on_random_field(hip_device):
    geom = O.build_geometry(96, 80, 60.0, 50.0)
    rng = np.random.default_rng(3)
    f = 0.4 + 0.2 * rng.normal(size=(80, 96))
    for model in ("std", "log"):
        p = O.Params(model=model, beta=0.7, q0=1.1, epsilon=-0.3, g=0.4)
        sim = O.Simulation(p, geom)
        sim.set_field(np.abs(f) if model == "log" else f)
        pl = _plan(geom, model=model, beta=0.7, q0=1.1, epsilon=-0.3, g=0.4)
        pl.set_field(sim.field)
        fe, mean, mn, mx = pl.reduce()
        assert fe == pytest.approx(sim.free_energy(), rel=1e-12)
        assert mean == pytest.approx(O.exact_mean(sim.field), rel=1e-14)
        assert (mn, mx) == (pytest.approx(sim.field.min(), abs=1e-13), pytest.approx(sim.field.max(), abs=1e-13))


def test_fp64_1024_etdrk2_matches_oracle(hip_device):
    """The bench workload itself (1024^2 seeded crystals, ETDRK2 dt = 0.5): 10 FP64 steps
    against the oracle, max|d psi| <= 1e-10."""
    from pfc_hip.plan import Plan
    n = 1024
    L = n * np.pi / 4
    geom = O.build_geometry(n, n, L, L)
    psi0 = F.seeded_crystals(geom, -0.3, -0.3)
    sim = O.Simulation(O.Params(model="std", epsilon=-0.3, g=0.0, dt=0.5, scheme="etdrk2"), geom)
    sim.set_field(psi0)
    sim.step(10)
    pl = Plan(Nx=n, Ny=n, Lx=L, Ly=L, epsilon=-0.3, g=0.0, dt=0.5, scheme="etdrk2")
    pl.set_field(psi0)
    pl.step(10)
    err = np.abs(pl.get_field().cpu().numpy() - sim.field).max()
    assert err <= 1e-10, err
    assert pl.free_energy() == pytest.approx(sim.free_energy(), rel=1e-12)


@pytest.mark.parametrize("precision", [64, 32])
def test_largest_grid_4096_matches_oracle(hip_device, precision):
    """4096 x 4096, the largest column length the column kernel holds (one
    4096-point column in LDS: 128 KiB FP64), stdPFC with g != 0, ETDRK2 dt =
    0.5 from seeded noise: 5 steps against the oracle.  FP64: max|d psi| <=
    1e-10 and free energy to 1e-12; FP32 bulk: max|d psi| <= 1e-5 and free
    energy to 1e-6."""
    import os
    n = 4096
    L = n * np.pi / 4
    geom = O.build_geometry(n, n, L, L)
    psi0 = F.uniform_noise((n, n), -0.3, 0.05, seed=41)
    p = O.Params(model="std", epsilon=-0.3, g=0.5, dt=0.5, scheme="etdrk2")
    sim = O.Simulation(p, geom, fft="scipy", workers=min(16, os.cpu_count() or 1))
    sim.set_field(psi0)
    sim.step(5)
    pl = _plan(geom, precision, epsilon=-0.3, g=0.5, dt=0.5, scheme="etdrk2").set_field(psi0)
    pl.step(5)
    err = float(np.abs(pl.get_field(dtype=None).cpu().numpy() - sim.field).max())
    assert err <= (1e-10 if precision == 64 else 1e-5), err
    assert pl.free_energy() == pytest.approx(sim.free_energy(), rel=1e-12 if precision == 64 else 1e-6)


def test_grid_beyond_the_column_kernel_is_refused(hip_device):
    """A column longer than the column kernel holds fails at creation with
    PFC_EUNSUPPORTED, before any allocation (no silent fallback)."""
    from pfc_hip import _lib
    n = 8192
    with pytest.raises(_lib.PfcError) as e:
        _plan(O.build_geometry(64, n, 16.0, n * 0.8), 64, epsilon=-0.3, g=0.0, dt=0.5)
    assert e.value.code == _lib.PfcError.UNSUPPORTED


@pytest.mark.parametrize("shape", [(128, 128), (96, 192), (256, 384), (64, 2048)])
@pytest.mark.parametrize("precision", [64, 32])
def test_fast_kernels_match_generic(hip_device, shape, precision, monkeypatch):
    """The compile-time-length kernels and the generic mixed-radix kernels compute the
    same step (FP64: <= 1e-13 after 20 ETDRK2 steps; FP32: <= 2e-5)."""
    from pfc_hip.plan import Plan
    ny, nx = shape
    geom = O.build_geometry(nx, ny, nx * np.pi / 4, ny * np.pi / 4)
    psi0 = F.uniform_noise(shape, -0.3, 0.05, seed=5)
    kw = dict(Nx=nx, Ny=ny, Lx=geom.Lx, Ly=geom.Ly, epsilon=-0.3, g=0.2, dt=0.5, scheme="etdrk2",
              precision=precision)
    fast = Plan(**kw).set_field(psi0).step(20)
    monkeypatch.setenv("PFC_HIP_GENERIC", "1")
    gen = Plan(**kw).set_field(psi0).step(20)
    d = np.abs(fast.get_field().cpu().numpy() - gen.get_field().cpu().numpy()).max()
    assert d <= (1e-13 if precision == 64 else 2e-5), d


def test_fp32_step_tolerance(hip_device):
    """FP32 bulk against the FP64 oracle, 128^2 ETD1 x 100 (r = psibar = -0.25, dt = 0.5):
    max|d psi| <= 1e-5, relative FE <= 1e-6, mean held exactly by the k = 0 pin (1e-7)."""
    from pfc_hip.plan import Plan
    geom, p, psi0 = _std128("etd1")
    sim = O.Simulation(p, geom)
    sim.set_field(psi0)
    sim.step(100)
    pl = Plan(Nx=128, Ny=128, Lx=geom.Lx, Ly=geom.Ly, epsilon=-0.25, g=0.0, dt=0.5, precision=32)
    pl.set_field(psi0)
    m0 = pl.mean()
    pl.step(100)
    err = np.abs(pl.get_field().cpu().numpy() - sim.field).max()
    fe, mean, _, _ = pl.reduce()
    assert err <= 1e-5, err
    assert abs(fe - sim.free_energy()) <= 1e-6 * abs(sim.free_energy())
    assert abs(mean - m0) <= 1e-7


# schedule variants per length (mirrors the library's table): {L: [(variant, points per lane)]}
SCHEDULES = {32: [(0, 8)], 48: [(0, 12)], 64: [(0, 8), (1, 4)], 96: [(0, 12)], 128: [(0, 8), (1, 4)],
             192: [(0, 12)], 256: [(0, 8), (1, 4), (2, 16)], 384: [(0, 12)], 512: [(0, 8), (1, 4), (2, 16)],
             768: [(0, 12)], 1024: [(0, 16), (1, 8), (2, 4)], 2048: [(0, 16), (1, 8)]}


def _shapes(L, lds_elems, cs):
    """(variant, lines) pairs the library accepts for a length-L line."""
    out = []
    for var, P in SCHEDULES[L]:
        tpl = L // P
        w = 1
        while w <= 256:
            if 64 <= tpl * w <= 256 and lds_elems(w) * cs <= 163840:
                out.append((var, w))
            w *= 2
    return out


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("shape", [(1024, 1024), (256, 512), (192, 384)])
def test_every_schedule_matches_generic(hip_device, shape, precision, monkeypatch):
    """Each (schedule variant, lines per block) the autotuner may pick computes the
    same ETDRK2 step as the generic kernels (FP64 <= 1e-13, FP32 <= 2e-5 after 4 steps)."""
    from pfc_hip.plan import Plan
    ny, nx = shape
    cs = 16 if precision == 64 else 8
    psi0 = F.uniform_noise(shape, -0.3, 0.05, seed=8)
    kw = dict(Nx=nx, Ny=ny, Lx=nx * np.pi / 4, Ly=ny * np.pi / 4, epsilon=-0.3, g=0.2, dt=0.5, scheme="etdrk2",
              precision=precision)
    monkeypatch.setenv("PFC_HIP_GENERIC", "1")
    ref = Plan(**kw).set_field(psi0).step(4).get_field().cpu().numpy()
    monkeypatch.delenv("PFC_HIP_GENERIC")
    lpad = lambda i: i + (i >> 4)  # LDS padding of the library's fast kernels
    rows = _shapes(nx // 2, lambda w: w * (lpad(nx // 2) + 2), cs)
    cols = _shapes(ny, lambda w: w * (lpad(ny - 1) + 2), cs)
    tol = 1e-13 if precision == 64 else 2e-5
    combos = [(r, cols[0]) for r in rows] + [(rows[0], c) for c in cols[1:]]
    for (rv, rl), (cv, cl) in combos:
        monkeypatch.setenv("PFC_HIP_SCHEDULE", f"{rv}:{rl}:{cv}:{cl}")
        pl = Plan(**kw).set_field(psi0).step(4)
        sched = pl.schedule()
        assert sched["row"] == (1, rv, dict(SCHEDULES[nx // 2])[rv], rl)
        assert sched["col"] == (1, cv, dict(SCHEDULES[ny])[cv], cl)
        d = np.abs(pl.get_field().cpu().numpy() - ref).max()
        assert d <= tol, (rv, rl, cv, cl, d)


@pytest.mark.parametrize("shape", [(1024, 1024), (256, 512)])
def test_capped_kernel_forms_are_bitwise_equal(hip_device, shape, monkeypatch):
    """The FP32 kernel forms held to three waves per SIMD (variant + 16, an
    autotuner candidate for schedules of up to 8 points per lane) compute the
    same ETDRK2 step as their uncapped twins, bit for bit: only the register
    allocation differs."""
    from pfc_hip.plan import Plan
    ny, nx = shape
    psi0 = F.uniform_noise(shape, -0.3, 0.05, seed=9)
    kw = dict(Nx=nx, Ny=ny, Lx=nx * np.pi / 4, Ly=ny * np.pi / 4, epsilon=-0.3, g=0.2, dt=0.5, scheme="etdrk2",
              precision=32)
    lpad = lambda i: i + (i >> 4)
    rows = [s for s in _shapes(nx // 2, lambda w: w * (lpad(nx // 2) + 2), 8) if dict(SCHEDULES[nx // 2])[s[0]] <= 8]
    cols = [s for s in _shapes(ny, lambda w: w * (lpad(ny - 1) + 2), 8) if dict(SCHEDULES[ny])[s[0]] <= 8]
    for (rv, rl), (cv, cl) in [(r, cols[0]) for r in rows] + [(rows[0], c) for c in cols[1:]]:
        fields = []
        for rcode, ccode in [(rv, cv), (rv + 16, cv + 16)]:
            monkeypatch.setenv("PFC_HIP_SCHEDULE", f"{rcode}:{rl}:{ccode}:{cl}")
            pl = Plan(**kw).set_field(psi0).step(6)
            assert pl.schedule()["row"][1] == rcode and pl.schedule()["col"][1] == ccode
            fields.append(pl.get_field().cpu().numpy())
        for k, f in enumerate(fields[1:]):
            assert np.array_equal(fields[0], f), (rv, rl, cv, cl, k)
