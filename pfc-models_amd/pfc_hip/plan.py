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


def make_params(*, model="std", Nx, Ny, Lx, Ly, Gamma=1.0, beta=1.0, q0=1.0, epsilon=1.0, g=1.0, dt=0.01,
                minPhi=1e-6, smallZTol=2e-5, M=0.0, precision=64, scheme="etd1", clamp_log=False, pin_mean=True):
    """`pfc_params` from keyword arguments (names of Build2DGeometry and the model builders)."""
    if scheme not in SCHEMES:
        raise ValueError(f"scheme must be one of {sorted(SCHEMES)}")
    if model not in MODELS:
        raise ValueError(f"model must be one of {sorted(MODELS)}")
    return _lib.PfcParams(
        model=MODELS[model], Nx=int(Nx), Ny=int(Ny), Lx=float(Lx), Ly=float(Ly),
        Gamma=float(Gamma), beta=float(beta), q0=float(q0), epsilon=float(epsilon), g=float(g),
        dt=float(dt), minPhi=float(minPhi), smallZTol=float(smallZTol), M=float(M),
        precision=int(precision), scheme=SCHEMES[scheme], clamp_log=int(bool(clamp_log)),
        pin_mean=int(bool(pin_mean)))


def slot_bytes(**kw):
    """Device bytes one instance of these parameters needs (state and tables)."""
    v = ctypes.c_longlong()
    p = make_params(**kw)
    _lib.check(_lib.load().pfc_slot_bytes(ctypes.byref(p), ctypes.byref(v)))
    return v.value


def _field_tensor(torch, x, Ny, Nx):
    if isinstance(x, torch.Tensor):
        t = x.to(device="cuda", dtype=torch.float64 if x.dtype == torch.float64 else torch.float32)
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(x, dtype=np.float64)), device="cuda")
    return t.reshape(Ny, Nx).contiguous()


class Plan:
    """One Ny x Nx instance resident in HBM: row-spectrum state, FP64-built
    ETD tables and reduction scratch, owned by the library."""

    def __init__(self, *, model="std", Nx, Ny, Lx, Ly, Gamma=1.0, beta=1.0, q0=1.0, epsilon=1.0, g=1.0,
                 dt=0.01, minPhi=1e-6, smallZTol=2e-5, M=0.0, precision=64, scheme="etd1",
                 clamp_log=False, pin_mean=True):
        self._lib = _lib.load()
        self.torch = _torch()
        p = make_params(model=model, Nx=Nx, Ny=Ny, Lx=Lx, Ly=Ly, Gamma=Gamma, beta=beta, q0=q0, epsilon=epsilon,
                        g=g, dt=dt, minPhi=minPhi, smallZTol=smallZTol, M=M, precision=precision, scheme=scheme,
                        clamp_log=clamp_log, pin_mean=pin_mean)
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.precision = int(precision)
        self.scheme = scheme
        self.model = model
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
        return _field_tensor(self.torch, x, self.Ny, self.Nx)

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

    def set_external_potential(self, V):
        """External potential V ((Ny, Nx) array; None removes it): the update
        gains P lap V_hat (pfc_set_external_potential)."""
        if V is None:
            _lib.check(self._lib.pfc_set_external_potential(self._h, None, 64, current_stream_handle()))
            return self
        t = self._device_field(V)
        dt = 64 if t.dtype == self.torch.float64 else 32
        _lib.check(self._lib.pfc_set_external_potential(self._h, ctypes.c_void_p(t.data_ptr()), dt,
                                                        current_stream_handle()))
        return self

    def diagnose(self, ky, kx):
        """Spectral diagnostics of the field (pfc_diagnose): (mean, m0, m1, m2,
        n_hat at the requested (ky, kx) as a complex array)."""
        return _unpack_diag(_diagnose(self._lib.pfc_diagnose, self._h, ky, kx, 1))[0]

    def step_profile(self, nsteps):
        """Each kernel kind of the current schedule, `nsteps` launches back to
        back on scratch copies of the state (the state is not advanced), timed
        with one HIP event pair per kind; returns {kind: (total_ms, calls)} for
        kinds 'row', 'col', 'col2'."""
        ms = (ctypes.c_double * 3)()
        calls = (ctypes.c_longlong * 3)()
        _lib.check(self._lib.pfc_step_profile(self._h, int(nsteps), current_stream_handle(), ms, calls))
        return {k: (ms[i], calls[i]) for i, k in enumerate(("row", "col", "col2"))}

    def reduce(self):
        """(free energy per site, mean, min, max) in FP64; synchronises."""
        _lib.check(self._lib.pfc_reduce(self._h, self._out, current_stream_handle()))
        return tuple(float(v) for v in self._out)

    def run_tracked(self, nsteps, every):
        """Advance `nsteps` steps, recording (free energy per site, mean, min,
        max) in FP64 after every `every` steps: an (nsteps // every, 4) array,
        synchronised once."""
        nrec = int(nsteps) // int(every) if int(every) > 0 else 0
        out = np.zeros((max(nrec, 1), 4), dtype=np.float64)
        _lib.check(self._lib.pfc_run_tracked(self._h, int(nsteps), int(every),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             current_stream_handle()))
        return out[:nrec]

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

    def set_auto_shift(self, on=True, floor_tol=0.1, ceil_tol=0.1):
        """Let the device refit the stabilising shift M before every step
        (pfc_set_auto_shift); on=False hands M back to set_shift; on=2 (the
        shift already on) changes the tolerances and keeps the current fit."""
        _lib.check(self._lib.pfc_set_auto_shift(self._h, 2 if on == 2 else int(bool(on)), float(floor_tol),
                                                float(ceil_tol), current_stream_handle()))
        return self

    def shift(self):
        """(M, lo, hi) of the device-managed shift (lo = hi = nan before the first fit)."""
        v = (ctypes.c_double * 3)()
        _lib.check(self._lib.pfc_get_shift(self._h, v, current_stream_handle()))
        return float(v[0]), float(v[1]), float(v[2])

    def status(self):
        v = ctypes.c_int()
        _lib.check(self._lib.pfc_get_status(self._h, ctypes.byref(v), current_stream_handle()))
        return v.value

    def schedule(self):
        """Kernel schedule in use: {'row': (fast, variant, P, lines), 'col': (...)}; variant + 16 is
        the FP32 kernel form held to three waves per SIMD; fast = -1 / -2: the resident stepper
        in its FFT / dense form."""
        info = (ctypes.c_int * 8)()
        _lib.check(self._lib.pfc_plan_schedule(self._h, info))
        return {"row": tuple(info[0:4]), "col": tuple(info[4:8])}

    def stores(self):
        """Output store kind of the fast kernels: 'nt' (non-temporal),
        'ordinary', or None when they do not both run."""
        v = ctypes.c_int()
        _lib.check(self._lib.pfc_plan_stores(self._h, ctypes.byref(v)))
        return {1: "nt", 0: "ordinary"}.get(v.value)

    def graph_stats(self):
        """(step-loop graphs captured, graph replays) so far."""
        v = (ctypes.c_longlong * 2)()
        _lib.check(self._lib.pfc_plan_graph_stats(self._h, v))
        return int(v[0]), int(v[1])

    def device_bytes(self):
        v = ctypes.c_longlong()
        _lib.check(self._lib.pfc_plan_bytes(self._h, ctypes.byref(v)))
        return v.value


class Batch:
    """`capacity` same-shape instances stepped by the same launches.

    Every pfc_params field is shared except epsilon and M (`set_slot`) and
    the box Lx, Ly and time step dt (`set_slot_box`), which are per slot.
    `step(n)` advances the active slots (initially all)
    together; `reduce()` returns an (n_active, 4) array of FP64
    (free energy per site, mean, min, max)."""

    def __init__(self, capacity, **kw):
        self._lib = _lib.load()
        self.torch = _torch()
        p = make_params(**kw)
        self.params = p
        self.capacity = int(capacity)
        self.Nx, self.Ny = p.Nx, p.Ny
        self.precision = p.precision
        h = ctypes.c_void_p()
        _lib.check(self._lib.pfc_batch_create(ctypes.byref(p), self.capacity, ctypes.byref(h)))
        self._h = h
        self._active = list(range(self.capacity))
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.torch.cuda.synchronize()
            _lib.check(self._lib.pfc_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_slot(self, slot, epsilon, M=0.0):
        _lib.check(self._lib.pfc_batch_set_slot(self._h, int(slot), float(epsilon), float(M),
                                                current_stream_handle()))
        return self

    def set_slot_box(self, slot, Lx, Ly, dt):
        """The slot's box and time step (its k-grids and ETD tables are rebuilt)."""
        _lib.check(self._lib.pfc_batch_set_slot_box(self._h, int(slot), float(Lx), float(Ly), float(dt),
                                                    current_stream_handle()))
        return self

    def set_field(self, slot, x):
        t = _field_tensor(self.torch, x, self.Ny, self.Nx)
        dt = 64 if t.dtype == self.torch.float64 else 32
        _lib.check(self._lib.pfc_batch_set_field(self._h, int(slot), ctypes.c_void_p(t.data_ptr()), dt,
                                                 current_stream_handle()))
        self._keep[int(slot)] = t  # must outlive the asynchronous upload
        return self

    def get_field(self, slot, out=None, dtype=None):
        torch = self.torch
        if out is None:
            out = torch.empty((self.Ny, self.Nx), device="cuda", dtype=dtype or torch.float64)
        dt = 64 if out.dtype == torch.float64 else 32
        _lib.check(self._lib.pfc_batch_get_field(self._h, int(slot), ctypes.c_void_p(out.data_ptr()), dt,
                                                 current_stream_handle()))
        return out

    def set_external_potential(self, slot, V):
        """The slot's external potential ((Ny, Nx) array; None removes it)."""
        if V is None:
            _lib.check(self._lib.pfc_batch_set_external_potential(self._h, int(slot), None, 64,
                                                                  current_stream_handle()))
            return self
        t = _field_tensor(self.torch, V, self.Ny, self.Nx)
        dt = 64 if t.dtype == self.torch.float64 else 32
        _lib.check(self._lib.pfc_batch_set_external_potential(self._h, int(slot), ctypes.c_void_p(t.data_ptr()), dt,
                                                              current_stream_handle()))
        return self

    def get_fields(self, dtype=None):
        """Fields of the active slots in active order, one device array (n_active, Ny, Nx)."""
        torch = self.torch
        out = torch.empty((len(self._active), self.Ny, self.Nx), device="cuda", dtype=dtype or torch.float64)
        if self._active:
            dt = 64 if out.dtype == torch.float64 else 32
            _lib.check(self._lib.pfc_batch_get_fields(self._h, ctypes.c_void_p(out.data_ptr()), dt,
                                                      current_stream_handle()))
        return out

    def diagnose(self, ky, kx):
        """pfc_batch_diagnose: per active slot (mean, m0, m1, m2, n_hat entries)."""
        return _unpack_diag(_diagnose(self._lib.pfc_batch_diagnose, self._h, ky, kx, len(self._active)))

    def set_active(self, slots):
        slots = [int(s) for s in slots]
        arr = (ctypes.c_int * max(1, len(slots)))(*slots)
        _lib.check(self._lib.pfc_batch_set_active(self._h, arr, len(slots), current_stream_handle()))
        self._active = slots
        return self

    @property
    def active(self):
        return list(self._active)

    def step(self, nsteps=1):
        _lib.check(self._lib.pfc_batch_step(self._h, int(nsteps), current_stream_handle()))
        return self

    def step_counts(self, counts):
        """counts[i] steps for the i-th active slot (resident-stepper batches)."""
        c = np.ascontiguousarray(np.asarray(counts, dtype=np.int64))
        if c.size != len(self._active):
            raise ValueError("one count per active slot")
        _lib.check(self._lib.pfc_batch_step_counts(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                                   current_stream_handle()))
        return self

    def reduce(self):
        n = len(self._active)
        out = (ctypes.c_double * max(4, 4 * n))()
        _lib.check(self._lib.pfc_batch_reduce(self._h, out, current_stream_handle()))
        return np.array(out[: 4 * n], dtype=np.float64).reshape(n, 4)

    def set_auto_shift(self, on=True, floor_tol=0.1, ceil_tol=0.1):
        """Device-managed stabilising shift for every slot (see Plan.set_auto_shift)."""
        _lib.check(self._lib.pfc_batch_set_auto_shift(self._h, 2 if on == 2 else int(bool(on)), float(floor_tol),
                                                      float(ceil_tol), current_stream_handle()))
        return self

    def shift(self, slot):
        """(M, lo, hi) of `slot`'s device-managed shift."""
        v = (ctypes.c_double * 3)()
        _lib.check(self._lib.pfc_batch_get_shift(self._h, int(slot), v, current_stream_handle()))
        return float(v[0]), float(v[1]), float(v[2])

    def run_tracked(self, nsteps, every):
        """Advance the active slots `nsteps` steps, recording each one's (free
        energy per site, mean, min, max) after every `every` steps: an
        (nsteps // every, n_active, 4) FP64 array, synchronised once."""
        nrec = int(nsteps) // int(every) if int(every) > 0 else 0
        n = len(self._active)
        out = np.zeros((max(nrec, 1), max(n, 1), 4), dtype=np.float64)
        _lib.check(self._lib.pfc_batch_run_tracked(self._h, int(nsteps), int(every),
                                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   current_stream_handle()))
        return out[:nrec, :n]

    def status(self):
        v = (ctypes.c_int * self.capacity)()
        _lib.check(self._lib.pfc_batch_get_status(self._h, v, current_stream_handle()))
        return np.array(v[:], dtype=np.int32)

    def graph_stats(self):
        """(step-loop graphs captured, graph replays) so far."""
        v = (ctypes.c_longlong * 2)()
        _lib.check(self._lib.pfc_batch_graph_stats(self._h, v))
        return int(v[0]), int(v[1])

    def device_bytes(self):
        v = ctypes.c_longlong()
        _lib.check(self._lib.pfc_batch_bytes(self._h, ctypes.byref(v)))
        return v.value


class Group:
    """Batches of different grid sizes advanced by shared launches.

    `step(n)` advances the active slots of every member batch with one row
    launch and one column launch per nonlinear evaluation for the whole group
    (pfc_group_step).  Members share precision and scheme and have Nx/2 and
    Ny in {32, 48, 64, 96, 128, 192, 256, 384, 512}; they stay usable on
    their own (set fields, reduce, read).  The group keeps its members alive."""

    def __init__(self, batches, flow=0):
        self._lib = _lib.load()
        self.torch = _torch()
        self.batches = list(batches)
        arr = (ctypes.c_void_p * len(self.batches))(*[b._h.value for b in self.batches])
        h = ctypes.c_void_p()
        _lib.check(self._lib.pfc_group_create(arr, len(self.batches), ctypes.byref(h)))
        self._h = h
        if flow:
            self.set_flow(flow)

    def set_flow(self, steps_per_launch):
        """Persistent form (pfc_group_set_flow): each register class steps in
        one launch per `steps_per_launch` steps over a work list of tiles with
        per-instance dependencies; 0 returns to the per-pass launches."""
        _lib.check(self._lib.pfc_group_set_flow(self._h, int(steps_per_launch)))
        return self

    def flow_stats(self):
        """(launches so far, blocks per class-0 launch, blocks per class-1 launch, steps per launch)."""
        v = (ctypes.c_longlong * 4)()
        _lib.check(self._lib.pfc_group_flow_stats(self._h, v))
        return tuple(int(x) for x in v)

    def step(self, nsteps=1):
        _lib.check(self._lib.pfc_group_step(self._h, int(nsteps), current_stream_handle()))
        return self

    def graph_stats(self):
        """(step-loop graphs captured, graph replays) so far."""
        v = (ctypes.c_longlong * 2)()
        _lib.check(self._lib.pfc_group_graph_stats(self._h, v))
        return int(v[0]), int(v[1])

    def chunks(self):
        """Number of cache-sized chunks the last step call ran the members in."""
        v = ctypes.c_int()
        _lib.check(self._lib.pfc_group_chunks(self._h, ctypes.byref(v)))
        return int(v.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.torch.cuda.synchronize()
            _lib.check(self._lib.pfc_group_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _diagnose(fn, h, ky, kx, n):
    ky = np.asarray(ky, dtype=np.int32).ravel()
    kx = np.asarray(kx, dtype=np.int32).ravel()
    if ky.shape != kx.shape:
        raise ValueError("ky and kx must have the same length")
    ne = ky.size
    out = np.zeros((max(n, 1), 4 + 2 * ne), dtype=np.float64)
    iptr = ctypes.POINTER(ctypes.c_int)
    _lib.check(fn(h, ky.ctypes.data_as(iptr), kx.ctypes.data_as(iptr), ne,
                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), current_stream_handle()))
    return out[:n]


def _unpack_diag(raw):
    """rows of pfc_diagnose output -> [(mean, m0, m1, m2, entries)]"""
    return [(r[0], r[1], r[2], r[3], r[4::2] + 1j * r[5::2]) for r in raw]


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
