"""Batched sweep against the plain FP64 one-at-a-time path (needs a GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_fp32_sweep_matches_plain_fp64(hip_device):
    """A 4 x 4 corner of the (r, psibar) sweep over all six grid sizes, 300 ETD1
    steps.  FP32-bulk batched free energies against the plain FP64 step:
    median relative error <= 1e-6 and max <= 5e-6 (the instances above 1e-6
    are mid-crystallisation, where the linear instability amplifies FP32
    round-off in the noise-seeded pattern); means to 1e-6 absolute.  The FP64
    batched run agrees with the plain FP64 run to 1e-12."""
    from pfc_hip import sweep as SW
    inst = SW.sweep_instances(n_r=4, n_psi=4, r_range=(-0.5, -0.2), psi_range=(-0.45, -0.15))
    assert len({i.N for i in inst}) >= 4
    plain = SW.run_plain(inst, 300)
    b32 = SW.run_batched(inst, 300, precision=32)
    b64 = SW.run_batched(inst, 300, precision=64)
    assert (plain.nonfinite == 0).all() and (b32.nonfinite == 0).all()
    rel32 = np.abs(b32.fe - plain.fe) / np.abs(plain.fe)
    rel64 = np.abs(b64.fe - plain.fe) / np.abs(plain.fe)
    assert np.median(rel32) <= 1e-6, rel32
    assert rel32.max() <= 5e-6, rel32
    assert rel64.max() <= 1e-12, rel64
    assert np.abs(b32.mean - plain.mean).max() <= 1e-6
    psibar = np.array([i.psibar for i in inst])
    assert np.abs(plain.mean - psibar).max() <= 0.01


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
