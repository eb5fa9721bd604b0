"""Batched sweep against the plain FP64 one-at-a-time path (needs a GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_fp32_sweep_matches_plain_fp64(hip_device):
    """Every 8th instance of the 256-instance sweep (all six grid sizes),
    5000 ETD1 steps.  FP32-bulk batched free energies against the plain FP64
    step: <= 1e-6 relative for every instance whose FP64 free energy moved by
    <= 1e-6 relative over its last 500 steps (relaxed), median <= 2e-7 over
    all; instances still mid-nucleation (the linear instability amplifies
    round-off there) <= 1e-5.  Means to 1e-7 absolute; FP64 batched equals
    plain FP64 to 1e-12."""
    from pfc_hip import sweep as SW
    inst = SW.sweep_instances()[::8]
    assert len({i.N for i in inst}) == 6
    plain = SW.run_plain(inst, 5000, checkpoint=500)
    b32 = SW.run_batched(inst, 5000, precision=32)
    b64 = SW.run_batched(inst, 5000, precision=64)
    assert (plain.nonfinite == 0).all() and (b32.nonfinite == 0).all()
    rel32 = np.abs(b32.fe - plain.fe) / np.abs(plain.fe)
    rel64 = np.abs(b64.fe - plain.fe) / np.abs(plain.fe)
    relaxed = np.abs(plain.fe - plain.fe_prev) <= 1e-6 * np.abs(plain.fe)
    print(f"sweep FP32: relaxed {relaxed.sum()}/{len(inst)} max {rel32[relaxed].max():.2e}; "
          f"still relaxing max {rel32[~relaxed].max() if (~relaxed).any() else 0:.2e}; median {np.median(rel32):.2e}",
          flush=True)
    assert relaxed.sum() >= 16
    assert rel32[relaxed].max() <= 1e-6, rel32
    assert np.median(rel32) <= 2e-7, rel32
    assert rel32.max() <= 1e-5, rel32
    assert rel64.max() <= 1e-12, rel64
    assert np.abs(b32.mean - plain.mean).max() <= 1e-7


@pytest.mark.parametrize("index", [40, 104])
def test_fp64_sweep_instance_against_the_oracle_5000_steps(hip_device, index):
    """Two 128^2 sweep instances still relaxing at 5000 steps (nucleation under
    way, where round-off is amplified most): the plain FP64 GPU step against
    the FP64 oracle at every 500 steps, relative free energy <= 1e-9, and the
    same structure-factor peaks at the end."""
    import os
    from oracle import pfc_oracle as O
    from pfc_hip import sweep as SW
    from pfc_hip.plan import Plan
    i = SW.sweep_instances()[index]
    assert i.N == 128
    geom = O.build_geometry(i.N, i.N, i.Lx, i.Ly)
    ref = O.Simulation(O.Params(model="std", epsilon=i.r, g=0.0, dt=0.5, pin_mean=None), geom, fft="scipy",
                       workers=min(16, os.cpu_count() or 1))
    f0 = SW.initial_field(i)
    ref.set_field(f0)
    pl = Plan(**SW._model_kw(i, 0.5, 64, "etd1")).set_field(f0)
    worst = 0.0
    for _ in range(10):
        ref.step(500)
        pl.step(500)
        fe = pl.reduce()[0]
        worst = max(worst, abs(fe - ref.free_energy()) / abs(ref.free_energy()))
    print(f"sweep instance {index}: FP64 GPU vs oracle max rel FE {worst:.2e}", flush=True)
    assert worst <= 1e-9
    assert O.structure_factor_peaks(pl.get_field().cpu().numpy(), 12) == O.structure_factor_peaks(ref.field, 12)


def test_coexistence_densities_fp32_vs_fp64(hip_device):
    """Liquid-crystal coexistence densities (common tangent of the analytic
    liquid and the relaxed crystals, 16 psibar values per r) from the FP32
    batched sweep against the FP64 batched sweep at three r values, 5000 ETD1
    steps: <= 1e-4 absolute wherever both find a tangent."""
    from pfc_hip import sweep as SW
    full = SW.sweep_instances()
    rs = sorted({i.r for i in full})
    pick = {rs[0], rs[4], rs[8]}
    inst = [i for i in full if i.r in pick]
    b32 = SW.run_batched(inst, 5000, precision=32)
    b64 = SW.run_batched(inst, 5000, precision=64)
    c32 = SW.coexistence_by_r(inst, b32.fe)
    c64 = SW.coexistence_by_r(inst, b64.fe)
    found = [r for r in pick if c64[r].status == "ok"]
    assert found, {r: c64[r].status for r in pick}
    for r in found:
        assert c32[r].status == "ok"
        assert abs(c32[r].phi_a - c64[r].phi_a) <= 1e-4, (r, c32[r], c64[r])
        assert abs(c32[r].phi_b - c64[r].phi_b) <= 1e-4, (r, c32[r], c64[r])


def test_grouped_sweep_matches_per_size_batches(hip_device):
    """The same sweep corner through one Group (shared launches over every
    size's tiles) and through per-size Batches on their own streams: FP64 free
    energies within 1e-12 relative (the per-size batches may run other
    autotuned radix schedules than the group's variant-0 ones), FP32 means to
    1e-6."""
    from pfc_hip import sweep as SW
    inst = SW.sweep_instances(n_r=4, n_psi=4, r_range=(-0.5, -0.2), psi_range=(-0.45, -0.15))
    for precision in (64, 32):
        a = SW.run_batched(inst, 200, precision=precision, grouped=True)
        b = SW.run_batched(inst, 200, precision=precision, grouped=False)
        assert (a.nonfinite == 0).all()
        if precision == 64:
            rel = np.abs(a.fe - b.fe) / np.abs(b.fe)
            assert rel.max() <= 1e-12, rel
        assert np.abs(a.mean - b.mean).max() <= 1e-6
