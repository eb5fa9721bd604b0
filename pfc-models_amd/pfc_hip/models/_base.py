"""Shared machinery of the `hip` model modules (stdpfc_hip, logpfc_hip).

The model API is the reference's (models/stdpfc/simulation.py:40-331,
models/logpfc/simulation.py:31-296): parameter builders, a Simulation with
TimeStep / CalculateFreeEnergy / Get_local_*, the `State` with the density
field, `t`, `mu`.  What differs, and why:

* The field lives on the device in the row-spectrum layout the fused kernels
  step; `State.n` / `State.phi` is a DeviceField proxy.  Assigning to it
  (`sim.State.n[...] = x`, x a NumPy array, a torch tensor or a scalar)
  uploads; reading it (`np.asarray`, `.get()`, indexing, arithmetic)
  downloads a host copy; `.mean()`, `.min()`, `.max()` are device FP64
  reductions.
* `mu` / `local_mu` are not by-products of the step (they do not feed the
  update); `_Batch2_CalcMu_Terms()` computes them on the device on request,
  as host arrays.
* New keywords: `precision` (64 or 32), `scheme` ('etd1' or 'etdrk2'),
  `pin_mean` (pin the k = 0 mode), and for logPFC `clamp_log` and `stabilise`.
  `run(nsteps)` advances many steps in one call.
"""
from __future__ import annotations

import numpy as np

from ..plan import Plan


def advance_time(t, dt, n):
    """The time a loop of `n` steps reaches from `t` by adding `dt` after every
    step, to the last bit (the additions run in the same order, as one
    np.add.accumulate), without the loop's per-step interpreter cost."""
    n = int(n)
    if n <= 0:
        return t
    a = np.empty(n + 1, dtype=np.float64)
    a[0] = t
    a[1:] = dt
    out = np.add.accumulate(a)[-1]
    return type(t)(out) if isinstance(t, np.floating) else float(out)

npFloatType = np.float64


class ModelParameters:
    """Attribute bag coercing int and float64 scalars to float64 and lists or
    tuples to float64 arrays (models/stdpfc/simulation.py:40-58)."""

    def __init__(self):
        return None

    def __setattr__(self, name, value):
        if isinstance(value, (int, npFloatType)) and not isinstance(value, bool):
            value = npFloatType(value)
        elif isinstance(value, (list, tuple)):
            value = np.array(value, dtype=npFloatType)
        super().__setattr__(name, value)


class DeviceField:
    """The density field of a hip Simulation, resident on the device."""

    __array_priority__ = 1000

    def __init__(self, sim):
        self._sim = sim

    # -- shape ------------------------------------------------------------
    @property
    def shape(self):
        return (self._sim.Geometry.Ny, self._sim.Geometry.Nx)

    @property
    def ndim(self):
        return 2

    @property
    def size(self):
        return self.shape[0] * self.shape[1]

    @property
    def dtype(self):
        return np.dtype(np.float64)

    def __len__(self):
        return self.shape[0]

    # -- transfers ----------------------------------------------------------
    def torch(self, dtype=None):
        """A fresh (Ny, Nx) device tensor holding the field."""
        return self._sim._plan.get_field(dtype=dtype)

    def get(self):
        """Host copy (NumPy float64), as CuPy's `.get()`."""
        return self._sim._plan.get_field().cpu().numpy()

    def copy(self):
        return self.get()

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def __getitem__(self, key):
        return self.get()[key]

    @staticmethod
    def _is_full(key):
        if key is Ellipsis or (isinstance(key, slice) and key == slice(None)):
            return True
        return (isinstance(key, tuple) and len(key) == 2
                and all(isinstance(k, slice) and k == slice(None) for k in key))

    def __setitem__(self, key, value):
        if isinstance(value, DeviceField):
            value = value.get()
        if self._is_full(key):
            if np.isscalar(value):
                value = np.full(self.shape, float(value))
            self._sim._upload(value)
        else:
            a = self.get()
            a[key] = np.asarray(value.cpu().numpy() if hasattr(value, "cpu") else value)
            self._sim._upload(a)

    # -- device reductions ----------------------------------------------------
    def mean(self, *args, **kwargs):
        if args or kwargs:
            return self.get().mean(*args, **kwargs)
        return npFloatType(self._sim._plan.reduce()[1])

    def min(self, *args, **kwargs):
        if args or kwargs:
            return self.get().min(*args, **kwargs)
        return npFloatType(self._sim._plan.reduce()[2])

    def max(self, *args, **kwargs):
        if args or kwargs:
            return self.get().max(*args, **kwargs)
        return npFloatType(self._sim._plan.reduce()[3])

    # -- host arithmetic on a copy ------------------------------------------
    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        inputs = tuple(i.get() if isinstance(i, DeviceField) else i for i in inputs)
        return getattr(ufunc, method)(*inputs, **kwargs)

    def _bin(op):
        def f(self, other):
            other = other.get() if isinstance(other, DeviceField) else other
            return op(self.get(), other)
        return f

    def _rbin(op):
        def f(self, other):
            return op(other, self.get())
        return f

    import operator as _op
    __add__ = _bin(_op.add)
    __radd__ = _rbin(_op.add)
    __sub__ = _bin(_op.sub)
    __rsub__ = _rbin(_op.sub)
    __mul__ = _bin(_op.mul)
    __rmul__ = _rbin(_op.mul)
    __truediv__ = _bin(_op.truediv)
    __rtruediv__ = _rbin(_op.truediv)
    __pow__ = _bin(_op.pow)
    del _op, _bin, _rbin

    def __neg__(self):
        return -self.get()

    def __abs__(self):
        return np.abs(self.get())

    def __repr__(self):
        g = self._sim.Geometry
        return f"DeviceField(shape=({g.Ny}, {g.Nx}), precision={self._sim.precision})"


class _SimulationState:
    """`State` of a hip Simulation: time, the device field and on-demand diagnostics."""

    def __init__(self, sim, field_name):
        self._sim = sim
        self._field_name = field_name
        self.t = npFloatType(0)
        self.mu = None
        self.local_mu = None
        setattr(self, field_name, DeviceField(sim))

    def __setattr__(self, name, value):
        object.__setattr__(self, name, value)
        # the reference reads the shift's window tolerances at every check
        # (logpfc_extv/simulation.py:375-405): a change reaches the device's
        # check at the next step, the current fit kept
        if name in ("phi_floor_tol", "phi_ceiling_tol"):
            sim = self.__dict__.get("_sim")
            if sim is not None and getattr(sim, "_device_shift", False):
                sim._plan.set_auto_shift(2, self.__dict__["phi_floor_tol"], self.__dict__["phi_ceiling_tol"])

    def __getattr__(self, name):
        # the stabilising shift when the device manages it (logPFC)
        if name in ("M", "phi_min", "phi_max"):
            sim = self.__dict__.get("_sim")
            if sim is not None and getattr(sim, "_device_shift", False):
                M, lo, hi = sim._plan.shift()
                v = {"M": M, "phi_min": lo, "phi_max": hi}[name]
                return v if name == "M" or v == v else None
        raise AttributeError(name)

    @property
    def kSpaceShape(self):
        g = self._sim.Geometry
        return (g.Ny, g.Nx // 2 + 1)

    @property
    def rSpaceShape(self):
        g = self._sim.Geometry
        return (g.Ny, g.Nx)

    def field_hat(self):
        """rfft2 of the field as a complex device tensor."""
        from ..plan import rfft2
        return rfft2(self._sim._plan.get_field())


def plan_kwargs(kind, modelParameters, geometry, *, precision=64, scheme="etd1", clamp_log=False, pin_mean=True):
    """The library parameters of one instance of model `kind` ('std' or 'log'):
    what a Simulation's plan and a batch slot are built from."""
    m = modelParameters
    return dict(model=kind, Nx=geometry.Nx, Ny=geometry.Ny, Lx=geometry.Lx, Ly=geometry.Ly, Gamma=float(m.Gamma),
                beta=float(m.beta), q0=float(m.q0), epsilon=float(m.epsilon), g=float(m.g), dt=float(m.dt),
                minPhi=float(getattr(m, "minPhi", 1e-6) or 1e-6), smallZTol=float(m.smallZTol), M=0.0,
                precision=int(precision), scheme=scheme, clamp_log=clamp_log, pin_mean=pin_mean)


class Simulation:
    """One instance on the device.  Built by BuildHipGPU_Simulation."""

    _model_kind = "std"
    _field_name = "n"

    def __init__(self, modelParameters, geometry, *, precision=64, scheme="etd1", pin_mean=True,
                 clamp_log=False, stabilise=False):
        self.Model = modelParameters
        self.Geometry = geometry
        self.precision = int(precision)
        self.scheme = scheme
        self.stabilise = bool(stabilise)
        self._plan = Plan(**plan_kwargs(self._model_kind, modelParameters, geometry, precision=self.precision,
                                        scheme=scheme, clamp_log=clamp_log, pin_mean=pin_mean))
        self.State = _SimulationState(self, self._field_name)
        self.mu = None
        # the field starts at zero, as the reference's preallocated buffers do
        self._upload(np.zeros((geometry.Ny, geometry.Nx)))

    # -- state ----------------------------------------------------------------
    @property
    def t(self):
        """Convenience alias for simulation time."""
        return self.State.t

    def _upload(self, value):
        self._plan.set_field(value)
        self.State.mu = self.State.local_mu = self.mu = None

    # -- stepping -------------------------------------------------------------
    def TimeStep(self):
        self.run(1)

    def run(self, nsteps, record_every=None):
        """Advance `nsteps` steps in one asynchronous call (t advances by nsteps*dt).

        With `record_every`, returns an (nsteps // record_every, 4) FP64 array of
        (free energy per site, mean, min, max) after every `record_every` steps,
        recorded on the device and synchronised once (the host-checked
        stabilised logPFC, stabilise="host", records block by block)."""
        nsteps = int(nsteps)
        rec = None
        if self.stabilise and getattr(self, "_host_shift", False):
            rows = []
            for i in range(nsteps):
                self.ValidatePhiRange()
                self._plan.step(1)
                self.State.t += self.Model.dt
                if record_every and (i + 1) % int(record_every) == 0:
                    rows.append(self._plan.reduce())
            if record_every:
                rec = np.array(rows, dtype=np.float64).reshape(-1, 4)
        else:
            if record_every:
                rec = self._plan.run_tracked(nsteps, int(record_every))
            else:
                self._plan.step(nsteps)
            # t accumulates exactly as a TimeStep loop does
            self.State.t = advance_time(self.State.t, self.Model.dt, nsteps)
        self.State.mu = self.State.local_mu = self.mu = None
        return rec

    # -- reference batches kept for callers that drive them directly ----------
    def _Batch1_CalcNonlinMu(self):
        """Nonlinear term and its transform: computed inside the device pass of
        _Batch2_CalcMu_Terms (nothing to keep between the two)."""
        return None

    def _Batch2_CalcMu_Terms(self):
        mu, lmu = self._plan.chemical_potential()
        self.State.mu = mu.cpu().numpy()
        self.State.local_mu = lmu.cpu().numpy()
        self.mu = self.State.mu

    def _Batch3_CalcNNext(self):
        raise NotImplementedError("the hip step fuses the update with its transforms; call TimeStep() or run(n)")

    # -- reductions -------------------------------------------------------------
    def CalculateFreeEnergy(self):
        """Box mean of the free-energy density, FP64 (compensated device sums)."""
        return npFloatType(self._plan.free_energy())

    def _local(self):
        return self._plan.local_field().cpu().numpy()

    def status(self):
        """Non-zero if a step met a non-finite value since the last call."""
        return self._plan.status()

    def close(self):
        self._plan.close()
