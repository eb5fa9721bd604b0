"""pfc_irfft2_noalloc: the irfft2 that allocates nothing (needs a GPU).

Called through the C ABI with ctypes on torch device buffers; its output is
checked against pfc_irfft2 (the allocating entry point with the same kernels)
and numpy.fft.irfft2, on two calls in a row, the second with a smaller input,
in FP64 and FP32.  Its roots are computed on the device, the allocating
path's on the host in long double, so the two may differ in a root's last bit:
FP64 outputs agree to 1e-14 of the field's scale, FP32 to 1e-6.  An input
beyond the static scratch is refused with PFC_EINVAL before any launch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, fn, X, Ny, Nx, torch):
    from pfc_hip.plan import current_stream_handle
    X = X.contiguous()
    batch = int(np.prod(X.shape[:-2])) if X.dim() > 2 else 1
    rdt = torch.float64 if X.dtype == torch.complex128 else torch.float32
    out = torch.full((*X.shape[:-2], Ny, Nx), float("nan"), dtype=rdt, device="cuda")
    rc = getattr(lib, fn)(ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(out.data_ptr()), batch, Ny, Nx,
                          64 if rdt == torch.float64 else 32, current_stream_handle())
    return rc, out


def _spectrum(rng, batch, Ny, Nx):
    x = rng.standard_normal((batch, Ny, Nx))
    return np.fft.rfft2(x), x


@pytest.mark.parametrize("dtype", [64, 32])
def test_noalloc_irfft2_matches_the_allocating_one(hip_device, dtype):
    import torch
    from pfc_hip import _lib
    lib = _lib.load()
    rng = np.random.default_rng(77)
    cdt = torch.complex128 if dtype == 64 else torch.complex64
    tol = 1e-14 if dtype == 64 else 1e-6
    # two calls in a row, the second smaller (the scratch keeps the first's data)
    for batch, Ny, Nx in [(3, 256, 384), (1, 96, 64)]:
        print(f"irfft2_noalloc dtype={dtype} batch={batch} Ny={Ny} Nx={Nx}", flush=True)
        Xh, x = _spectrum(rng, batch, Ny, Nx)
        X = torch.as_tensor(Xh, dtype=cdt, device="cuda")
        X0 = X.clone()
        rc_a, ref = _call(lib, "pfc_irfft2", X, Ny, Nx, torch)
        rc_b, got = _call(lib, "pfc_irfft2_noalloc", X, Ny, Nx, torch)
        torch.cuda.synchronize()
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, _lib.load().pfc_last_error())
        assert torch.equal(X, X0)  # the input is not modified
        g, r = got.cpu().numpy().astype(np.float64), ref.cpu().numpy().astype(np.float64)
        scale = float(np.abs(x).max())
        assert np.isfinite(g).all()
        assert float(np.abs(g - r).max()) <= tol * scale
        assert float(np.abs(g - x).max()) <= (1e-13 if dtype == 64 else 2e-5) * scale


def test_noalloc_irfft2_refuses_beyond_its_scratch(hip_device):
    import torch
    from pfc_hip import _lib
    lib = _lib.load()
    # 2 x 2048 x 1025 complex128 = 67 MB > 64 MiB: refused before any launch
    X = torch.zeros((2, 2048, 1025), dtype=torch.complex128, device="cuda")
    rc, out = _call(lib, "pfc_irfft2_noalloc", X, 2048, 2048, torch)
    assert rc == _lib.PfcError.INVALID
    assert b"static scratch" in lib.pfc_last_error()
    assert torch.isnan(out).all()
