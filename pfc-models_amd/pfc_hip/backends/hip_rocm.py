"""HIP / ROCm (AMD MI355X) backend.

Sibling of the reference's backend leaf modules (backends/cupy_cuda.py:19-32,
backends/numpy_fftw.py:17-28) with the fuller contract its stub documents
(backends/torch_mps.py:3-14):

    xp          array module (torch; device tensors on the HIP device)
    fftn/ifftn  not provided: the models need only the real 2D transforms
    rfftn/irfftn  2D real transforms over the last two axes (libpfc_hip.so)
    to_numpy(a) host copy as a numpy array
    available() -> bool
    name        str

Selection is explicit: nothing here falls back to the CPU.  `rfftn`/`irfftn`
match numpy.fft.rfft2 / irfft2 semantics (unnormalised forward, 1/N inverse;
the inverse drops the imaginary parts of the kx = 0 and Nyquist columns after
the column inverse, as numpy does).
"""
from __future__ import annotations

import numpy as np

npFloatType = np.float64
npComplexType = np.complex128

name = "hip"


def _torch():
    import torch
    return torch


try:
    import torch as xp  # noqa: F401  (device arrays are torch tensors)
    hipFloatType = xp.float64
    hipComplexType = xp.complex128
    hipFloat32Type = xp.float32
    hipComplex32Type = xp.complex64
except Exception:  # pragma: no cover - torch is part of this image
    xp = None


def available() -> bool:
    """True when a HIP device is visible and libpfc_hip.so loads."""
    try:
        torch = _torch()
        if not torch.cuda.is_available() or torch.version.hip is None:
            return False
        from .. import _lib
        _lib.load()
        return True
    except Exception:
        return False


def rfftn(a, axes=(-2, -1)):
    """numpy.fft.rfft2 of a real device tensor (float64 or float32) over the last two axes."""
    if tuple(axes) not in ((-2, -1), (0, 1)) or (tuple(axes) == (0, 1) and a.dim() != 2):
        raise ValueError("the hip backend transforms the last two axes only")
    from ..plan import rfft2
    return rfft2(a)


def irfftn(a, s, axes=(-2, -1)):
    """numpy.fft.irfft2(a, s) of a complex device tensor over the last two axes."""
    if tuple(axes) not in ((-2, -1), (0, 1)) or (tuple(axes) == (0, 1) and a.dim() != 2):
        raise ValueError("the hip backend transforms the last two axes only")
    from ..plan import irfft2
    return irfft2(a, s)


def to_numpy(a):
    """Host copy as a numpy array."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)
