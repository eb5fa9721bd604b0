"""Python handle of one device instance (a `pfc_plan` of libpfc_hip.so).

torch is used only for device memory and the current HIP stream; all compute
runs in the library's kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

SCHEMES = {"etd1": _lib.PFC_SCHEME_ETD1, "etdrk2": _lib.PFC_SCHEME_ETDRK2}
MODELS = {"std": _lib.PFC_MODEL_STD, "log": _lib.PFC_MODEL_LOG}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pfc_hip needs a HIP device (torch.cuda.is_available() is False)")
    return torch


def current_stream_handle():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Plan:
    """One Ny x Nx instance resident in HBM: row-spectrum state, FP64-built
    ETD tables and reduction scratch, owned by the library."""

    def __init__(self, *, model="std", Nx, Ny, Lx, Ly, Gamma=1.0, beta=1.0, q0=1.0, epsilon=1.0, g=1.0,
                 dt=0.01, minPhi=1e-6, smallZTol=2e-5, M=0.0, precision=64, scheme="etd1",
                 clamp_log=False, pin_mean=True):
        self._lib = _lib.load()
        self.torch = _torch()
        if scheme not in SCHEMES:
            raise ValueError(f"scheme must be one of {sorted(SCHEMES)}")
        if model not in MODELS:
            raise ValueError(f"model must be one of {sorted(MODELS)}")
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.precision = int(precision)
        self.scheme = scheme
        self.model = model
        p = _lib.PfcParams(
            model=MODELS[model], Nx=self.Nx, Ny=self.Ny, Lx=float(Lx), Ly=float(Ly),
            Gamma=float(Gamma), beta=float(beta), q0=float(q0), epsilon=float(epsilon), g=float(g),
            dt=float(dt), minPhi=float(minPhi), smallZTol=float(smallZTol), M=float(M),
            precision=self.precision, scheme=SCHEMES[scheme], clamp_log=int(bool(clamp_log)),
            pin_mean=int(bool(pin_mean)))
        self.params = p
        h = ctypes.c_void_p()
        _lib.check(self._lib.pfc_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self._out = (ctypes.c_double * 4)()

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.torch.cuda.synchronize()
            _lib.check(self._lib.pfc_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dtype(self):
        return self.torch.float64 if self.precision == 64 else self.torch.float32

    # -- field I/O ----------------------------------------------------------
    def _device_field(self, x):
        torch = self.torch
        if isinstance(x, torch.Tensor):
            t = x.to(device="cuda", dtype=torch.float64 if x.dtype == torch.float64 else torch.float32)
        else:
            t = torch.as_tensor(np.ascontiguousarray(np.asarray(x, dtype=np.float64)), device="cuda")
        t = t.reshape(self.Ny, self.Nx).contiguous()
        return t

    def set_field(self, x):
        t = self._device_field(x)
        dt = 64 if t.dtype == self.torch.float64 else 32
        _lib.check(self._lib.pfc_set_field(self._h, ctypes.c_void_p(t.data_ptr()), dt, current_stream_handle()))
        self._keep = t  # the copy must outlive the asynchronous upload kernel
        return self

    def get_field(self, out=None, dtype=None):
        torch = self.torch
        if out is None:
            out = torch.empty((self.Ny, self.Nx), device="cuda", dtype=dtype or torch.float64)
        if not (out.is_contiguous() and out.shape == (self.Ny, self.Nx)):
            raise ValueError("out must be a contiguous (Ny, Nx) device tensor")
        dt = 64 if out.dtype == torch.float64 else 32
        _lib.check(self._lib.pfc_get_field(self._h, ctypes.c_void_p(out.data_ptr()), dt, current_stream_handle()))
        return out

    # -- stepping and reductions ------------------------------------------
    def step(self, nsteps=1):
        _lib.check(self._lib.pfc_step(self._h, int(nsteps), current_stream_handle()))
        return self

    def step_profile(self, nsteps):
        """Step with HIP events around every launch; returns {kind: (total_ms, calls)}
        for kinds 'row', 'col', 'col2'."""
        ms = (ctypes.c_double * 3)()
        calls = (ctypes.c_longlong * 3)()
        _lib.check(self._lib.pfc_step_profile(self._h, int(nsteps), current_stream_handle(), ms, calls))
        return {k: (ms[i], calls[i]) for i, k in enumerate(("row", "col", "col2"))}

    def reduce(self):
        """(free energy per site, mean, min, max) in FP64; synchronises."""
        _lib.check(self._lib.pfc_reduce(self._h, self._out, current_stream_handle()))
        return tuple(float(v) for v in self._out)

    def free_energy(self):
        return self.reduce()[0]

    def mean(self):
        return self.reduce()[1]

    def chemical_potential(self, dtype=None):
        """(mu, local_mu) of the current field as (Ny, Nx) device tensors."""
        torch = self.torch
        dt = dtype or torch.float64
        mu = torch.empty((self.Ny, self.Nx), device="cuda", dtype=dt)
        lmu = torch.empty_like(mu)
        code = 64 if dt == torch.float64 else 32
        _lib.check(self._lib.pfc_chemical_potential(self._h, ctypes.c_void_p(mu.data_ptr()),
                                                    ctypes.c_void_p(lmu.data_ptr()), code, current_stream_handle()))
        return mu, lmu

    def local_field(self, dtype=None):
        """Gaussian-smoothed field irfft2(exp(-k^2 (2 pi)^2 / 2) n_hat) as a device tensor."""
        torch = self.torch
        dt = dtype or torch.float64
        out = torch.empty((self.Ny, self.Nx), device="cuda", dtype=dt)
        code = 64 if dt == torch.float64 else 32
        _lib.check(self._lib.pfc_local_field(self._h, ctypes.c_void_p(out.data_ptr()), code,
                                             current_stream_handle()))
        return out

    def set_shift(self, M):
        _lib.check(self._lib.pfc_set_shift(self._h, float(M)))

    def status(self):
        v = ctypes.c_int()
        _lib.check(self._lib.pfc_get_status(self._h, ctypes.byref(v), current_stream_handle()))
        return v.value

    def schedule(self):
        """Kernel schedule in use: {'row': (fast, variant, P, lines), 'col': (...)}."""
        info = (ctypes.c_int * 8)()
        _lib.check(self._lib.pfc_plan_schedule(self._h, info))
        return {"row": tuple(info[0:4]), "col": tuple(info[4:8])}

    def device_bytes(self):
        v = ctypes.c_longlong()
        _lib.check(self._lib.pfc_plan_bytes(self._h, ctypes.byref(v)))
        return v.value


def rfft2(x):
    """numpy.fft.rfft2 over the last two axes of a real device tensor (float64/float32)."""
    torch = _torch()
    lib = _lib.load()
    x = x.contiguous()
    Ny, Nx = x.shape[-2:]
    batch = int(np.prod(x.shape[:-2])) if x.dim() > 2 else 1
    cdt = torch.complex128 if x.dtype == torch.float64 else torch.complex64
    out = torch.empty((*x.shape[:-2], Ny, Nx // 2 + 1), dtype=cdt, device=x.device)
    dt = 64 if x.dtype == torch.float64 else 32
    _lib.check(lib.pfc_rfft2(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), batch, Ny, Nx, dt,
                             current_stream_handle()))
    return out


def irfft2(X, s):
    """numpy.fft.irfft2(X, s) over the last two axes of a complex device tensor."""
    torch = _torch()
    lib = _lib.load()
    X = X.contiguous()
    Ny, Nx = int(s[0]), int(s[1])
    if X.shape[-2:] != (Ny, Nx // 2 + 1):
        raise ValueError(f"spectrum shape {tuple(X.shape[-2:])} does not match s={s}")
    batch = int(np.prod(X.shape[:-2])) if X.dim() > 2 else 1
    rdt = torch.float64 if X.dtype == torch.complex128 else torch.float32
    out = torch.empty((*X.shape[:-2], Ny, Nx), dtype=rdt, device=X.device)
    dt = 64 if rdt == torch.float64 else 32
    _lib.check(lib.pfc_irfft2(ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(out.data_ptr()), batch, Ny, Nx, dt,
                              current_stream_handle()))
    return out
