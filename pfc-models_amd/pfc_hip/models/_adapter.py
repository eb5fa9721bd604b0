"""Shared machinery of the phase-diagram adapters (stdpfc_hip, logpfc_hip).

The adapter contract is the reference's (phase_diagram/q0star.py:5-20; adapters
models/stdpfc_cpu/phase_diagram.py, models/logpfc_cpu/phase_diagram.py):
structures as R2C (row, col) index sets of a rectangular unit cell of the
triangular lattice, a box scaled by the lattice scale q00, a 3-mode ansatz and
a 3-mode fitness.  Everything here runs on the host in FP64 on (Ny, Nx)
arrays; the simulation itself is the hip Simulation.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..discretization import Build2DGeometry
from .. import phase_diagram as _pd


@dataclass(frozen=True)
class Structure:
    """One target lattice: `kvectors` = (1-mode, 2-mode, 3-mode) sets of R2C
    (row, col) = (ky, kx) indices for one rectangular cell; `amplitude` is the
    ansatz scale (absolute for stdPFC, in units of phibar for logPFC);
    `maxima_per_cell` the expected density maxima per cell (None: ridges)."""
    name: str
    kvectors: tuple
    amplitudeSign: int
    amplitude: float
    maxima_per_cell: int | None


# the three R2C 3-mode sets of the rectangular cell (models/stdpfc_cpu/phase_diagram.py:55-92)
HEX_SETS = (((1, 1), (2, 0), (-2, 0), (-1, 1)), ((0, 2), (3, 1), (-3, 1)), ((2, 2), (4, 0), (-4, 0), (-2, 2)))
STRIPE_SETS = (((2, 0), (-2, 0)), ((4, 0), (-4, 0)), ((6, 0), (-6, 0)))

PARAM_KEYS = ('beta', 'g', 'v0', 'Gamma', 'q0')
CELL_KEYS = ('uNx', 'uNy', 'Mx', 'My')
ANSATZ_KEYS = ('mode', 'A1', 'A2', 'A3')


def validate_params(slug, params):
    """beta, g required; v0 must be 1 (the quartic coefficient of both models);
    Gamma, q0 default 1."""
    p = dict(params)
    unknown = sorted(set(p) - set(PARAM_KEYS))
    if unknown:
        raise ValueError(f'{slug}: unknown model parameters {unknown}; allowed {list(PARAM_KEYS)}')
    missing = [k for k in ('beta', 'g') if k not in p]
    if missing:
        raise ValueError(f'{slug}: missing model parameters {missing}')
    v0 = float(p.get('v0', 1.0))
    if v0 != 1.0:
        raise ValueError(f'{slug}: v0={v0} not supported; the quartic coefficient is 1 in this model')
    return {'beta': float(p['beta']), 'g': float(p['g']), 'v0': 1.0, 'Gamma': float(p.get('Gamma', 1.0)),
            'q0': float(p.get('q0', 1.0))}


def validate_cell(slug, cell):
    """uNx, uNy: grid points per rectangular unit cell; Mx, My: cells per box side."""
    c = dict(cell)
    if sorted(c) != sorted(CELL_KEYS):
        raise ValueError(f'{slug}: cell needs exactly {list(CELL_KEYS)}, got {sorted(c)}')
    out = {}
    for k in CELL_KEYS:
        v = c[k]
        if int(v) != v or int(v) < 1:
            raise ValueError(f'{slug}: cell {k}={v!r} must be a positive integer')
        out[k] = int(v)
    return out


def validate_ansatz(slug, ansatz):
    """mode (a label such as 'L2H'); A1, A2, A3: 1-, 2-, 3-mode amplitudes in
    units of the structure's ansatz scale."""
    a = dict(ansatz)
    if sorted(a) != sorted(ANSATZ_KEYS):
        raise ValueError(f'{slug}: ansatz needs exactly {list(ANSATZ_KEYS)}, got {sorted(a)}')
    if not isinstance(a['mode'], str):
        raise ValueError(f'{slug}: ansatz mode must be a string label')
    return {'mode': a['mode'], 'A1': float(a['A1']), 'A2': float(a['A2']), 'A3': float(a['A3'])}


def mode_indices(structure, cell):
    """The 1-, 2-, 3-mode (rows, cols) index arrays scaled to the Mx x My supercell."""
    out = []
    for pos in structure.kvectors:
        rows, cols = np.array(pos).T
        out.append((rows * cell['My'], cols * cell['Mx']))
    return out


def grid_shape(cell):
    return cell['uNy'] * cell['My'], cell['uNx'] * cell['Mx']


def box(cell, q00):
    """Rectangular cell of the triangular lattice with its first shell at q00:
    a = 4 pi / (sqrt 3 q00), h = 4 pi / q00; box Mx a x My h."""
    Ny, Nx = grid_shape(cell)
    a = 4 * np.pi / (3 ** 0.5 * q00)
    h = 4 * np.pi / q00
    return Build2DGeometry(Nx=Nx, Ny=Ny, Lx=cell['Mx'] * a, Ly=cell['My'] * h)


def three_mode(structure, cell, mean, A, B, C, nonnegative=False):
    """Field with the given mean and 1-, 2-, 3-mode amplitudes (irfft2 of a
    spectrum set at the structure's index sets); `nonnegative` rescales the
    modulation about the mean so that the minimum is zero when it would be
    negative (the logPFC ansatz)."""
    Ny, Nx = grid_shape(cell)
    size = Nx * Ny
    (r1, c1), (r2, c2), (r3, c3) = mode_indices(structure, cell)
    spec = np.zeros((Ny, Nx // 2 + 1), dtype=complex)
    spec[0, 0] = mean * size
    spec[r1, c1] = A * size
    spec[r2, c2] = B * size
    spec[r3, c3] = C * size
    field = np.fft.irfftn(spec, s=(Ny, Nx), axes=(0, 1)).real
    if nonnegative:
        lo = field.min()
        if lo < 0:
            field = mean + (field - mean) * mean / (mean - lo)
    return field


def fitness(field, structure, cell, relative=False, nonnegative=False):
    """1 - mean((field - field_3mode)^2) and the amplitudes (A, B, C): mean
    |rfft2(field)| over the 1-, 2-, 3-mode sets per site (divided by the mean
    when `relative`), field_3mode rebuilt from them."""
    Ny, Nx = grid_shape(cell)
    size = Nx * Ny
    spec = np.fft.rfft2(field)
    mean = spec.flat[0].real / size
    (r1, c1), (r2, c2), (r3, c3) = mode_indices(structure, cell)
    s = structure.amplitudeSign
    amps = []
    for r, c in ((r1, c1), (r2, c2), (r3, c3)):
        a = s * np.nanmean(np.abs(spec[r, c])) / size
        amps.append(a / mean if relative else a)
    A, B, C = amps
    k = mean if relative else 1.0
    rec = three_mode(structure, cell, mean, A * k, B * k, C * k, nonnegative=nonnegative)
    return 1 - float(np.abs(np.power(field - rec, 2).mean())), np.abs(A), np.abs(B), np.abs(C)


class SpectralFit:
    """`fitness` computed from a field's spectral diagnostics (pfc_diagnose:
    mean, m0 = mean of n^2, and n_hat = rfft2(n) at the structure's mode
    entries) instead of the field: amplitudes = sign * mean |n_hat| / N over
    each set (over the mean when relative), the 3-mode field rebuilt from
    them (rescaled to a zero minimum when nonnegative and negative), and by
    Parseval over the full spectrum
        mean((n - n3)^2) = m0 - mean^2 + sum_e w_e (|n_hat_e - t_e|^2 - |n_hat_e|^2) / N^2,
    t_e = the rebuilt field's real coefficient N a_set(e), w_e = 1 on the
    kx = 0 and Nx/2 columns, else 2.  Vectorised over many fields."""

    def __init__(self, structure, cell, relative=False, nonnegative=False):
        Ny, Nx = grid_shape(cell)
        self.N, self.relative, self.nonnegative = Nx * Ny, relative, nonnegative
        self.sign = structure.amplitudeSign
        ky, kx, which = [], [], []
        for i, (rows, cols) in enumerate(mode_indices(structure, cell)):
            ky += [int(r) % Ny for r in rows]
            kx += [int(c) for c in cols]
            which += [i] * len(rows)
        self.ky, self.kx = np.array(ky, dtype=np.int32), np.array(kx, dtype=np.int32)
        self.which = np.array(which)
        self.weight = np.where((self.kx == 0) | (2 * self.kx == Nx), 1.0, 2.0)
        self.nsets = len(structure.kvectors)
        if nonnegative:  # unit-amplitude field of each set, as three_mode builds it
            self.basis = np.empty((self.nsets, Ny, Nx))
            for i in range(self.nsets):
                spec = np.zeros((Ny, Nx // 2 + 1), dtype=complex)
                sel = self.which == i
                spec[self.ky[sel], self.kx[sel]] = self.N
                self.basis[i] = np.fft.irfftn(spec, s=(Ny, Nx), axes=(0, 1))

    def assemble(self, rows):
        """[(mean, m0, m1, m2, n_hat entries)] -> (fitness, amplitudes) arrays."""
        mean = np.array([r[0] for r in rows])
        m0 = np.array([r[1] for r in rows])
        ent = np.array([r[4] for r in rows]).reshape(len(rows), -1)
        mag = np.abs(ent)
        amps = np.stack([self.sign * mag[:, self.which == i].mean(axis=1) / self.N for i in range(self.nsets)], 1)
        if self.relative:
            amps = amps / mean[:, None]
        a = amps * (mean[:, None] if self.relative else 1.0)  # what three_mode receives
        if self.nonnegative:
            n3 = mean[:, None, None] + np.einsum('fs,syx->fyx', a, self.basis)
            lo = n3.min(axis=(1, 2))
            scale = np.where(lo < 0, mean / np.where(lo < 0, mean - lo, 1.0), 1.0)
            a = a * scale[:, None]
        t = a[:, self.which] * self.N
        extra = (self.weight * (np.abs(ent - t) ** 2 - mag ** 2)).sum(axis=1) / self.N ** 2
        resid = m0 - mean ** 2 + extra
        return 1.0 - np.abs(resid), np.abs(amps)


_FITS = {}


def spectral_fit(structure, cell, relative=False, nonnegative=False):
    """A cached SpectralFit per (structure, cell, relative, nonnegative)."""
    key = (structure, tuple(sorted(cell.items())), relative, nonnegative)
    if key not in _FITS:
        _FITS[key] = SpectralFit(structure, cell, relative, nonnegative)
    return _FITS[key]


def maxima_per_cell(field, cell):
    """Strict local maxima on the periodic grid (8 neighbours), per rectangular cell."""
    is_max = np.ones(field.shape, dtype=bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy or dx:
                is_max &= field > np.roll(field, (dy, dx), axis=(0, 1))
    return int(is_max.sum()) / (cell['Mx'] * cell['My'])


def diagnostics(field, structure, cell, relative=False, nonnegative=False):
    f, A, B, C = fitness(field, structure, cell, relative=relative, nonnegative=nonnegative)
    return _pd.Diagnostics(fitness=f, amplitudes=(A, B, C), details={'maxima_per_cell': maxima_per_cell(field, cell)})


class SlotBatch:
    """Driver evaluations of one model in the slots of a device batch
    (pfc_hip.plan.Batch): per slot the temperature (epsilon), the box of its
    q00, its dt and its field; the rest of the model is shared.  A slot is
    the batched counterpart of the Simulation `build` makes: the same library
    parameters (`plan_kwargs`), the same k-grids and tables, the same field
    transfers."""

    def __init__(self, kind, model_params, cell, capacity, fit, precision=64):
        from ..plan import Batch
        from ._base import plan_kwargs
        self.cell = cell
        self._fit = fit  # (structure, cell) -> SpectralFit
        self.batch = Batch(capacity, **plan_kwargs(kind, model_params, box(cell, 1.0), precision=precision))
        self._active = None

    def place(self, slot, temp, meanDensity, q00, dt, state, t=0.0):
        """Slot `slot` at temperature `temp` in the box of q00 with time step dt
        and field `state` (which carries the mean density); the models are
        autonomous, so the time `t` changes nothing here."""
        g = box(self.cell, q00)
        self.batch.set_slot(slot, temp)
        self.batch.set_slot_box(slot, g.Lx, g.Ly, dt)
        self.set_state(slot, state)

    def set_state(self, slot, state):
        self.batch.set_field(slot, np.asarray(state, dtype=np.float64))

    def state(self, slot):
        return self.batch.get_field(slot).cpu().numpy().astype(np.float64)

    def set_active(self, slots):
        slots = list(slots)
        if slots != self._active:
            self.batch.set_active(slots)
            self._active = slots

    def step(self, n):
        self.batch.step(n)

    def step_each(self, counts):
        """counts[i] steps for the i-th active slot; False if this batch cannot
        (grids beyond the resident stepper), and nothing is stepped then."""
        from .._lib import PfcError
        try:
            self.batch.step_counts(counts)
        except PfcError as e:
            if e.code == PfcError.UNSUPPORTED:
                return False
            raise
        return True

    def free_energy(self):
        """FP64 free energy per site of each active slot, in active order."""
        return self.batch.reduce()[:, 0].copy()

    def states(self):
        """Host fields of the active slots, in active order (one device call)."""
        return list(self.batch.get_fields().cpu().numpy())

    def diagnose_active(self, structure, cell):
        """The adapter's diagnostics of every active slot, in active order,
        computed on the device (details left empty)."""
        fit = self._fit(structure, cell)
        fitness, amps = fit.assemble(self.batch.diagnose(fit.ky, fit.kx))
        return [_pd.Diagnostics(fitness=float(f), amplitudes=tuple(float(x) for x in a)) for f, a in zip(fitness, amps)]

    def close(self):
        self.batch.close()


# -- frozen field (pfc_hip.phase_diagram.frozen_q) ---------------------------------
def k2_ref(cell):
    """|k|^2 on the full FFT grid of the q00 = 1 box, shape (Ny, Nx)."""
    Ny, Nx = grid_shape(cell)
    g = box(cell, 1.0)
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=g.Lx / Nx)
    ky = 2 * np.pi * np.fft.fftfreq(Ny, d=g.Ly / Ny)
    return kx[None, :] ** 2 + ky[:, None] ** 2


class FrozenField:
    """f(q) of one field frozen on the grid while the box scale q00 = q
    changes (every wavevector is q k_ref):
        f(q) = <local> + beta/2 (q0^4 A0 - 2 q0^2 q^2 A1 + q^4 A2),
        A_n = sum_k |phi_hat_k / N|^2 |k_ref|^(2n)  (full spectrum),
    so q_opt = q0 sqrt(A1 / A2), the one minimum when A2 > 0
    (models/logpfc_cpu/phase_diagram.py:475-500)."""

    def __init__(self, state, beta, q0, local_mean, cell):
        phi = np.asarray(state, dtype=np.float64)
        power = np.abs(np.fft.fft2(phi)) ** 2 / phi.size ** 2
        k2 = k2_ref(cell)
        self.A = np.array([power.sum(), (power * k2).sum(), (power * k2 * k2).sum()])
        self.local = float(local_mean)
        self.beta, self.q0 = float(beta), float(q0)

    def f(self, q):
        s = np.asarray(q, dtype=float) ** 2
        A0, A1, A2 = self.A
        q2 = self.q0 ** 2
        return self.local + self.beta / 2 * (q2 * q2 * A0 - 2 * q2 * s * A1 + s * s * A2)

    def q_opt(self):
        return self.q0 * np.sqrt(self.A[1] / self.A[2])


PEAK_SPREAD_MAX = 0.20
SUBQ_RATIO_MAX = 1e-3


def defect_diagnostics(state, cell):
    """Record-only defect screen (models/logpfc_cpu/phase_diagram.py:508-528):
    peak_spread = (max - min) / mean of the heights above the mean of the
    periodic 8-neighbour local maxima (>= every neighbour); subq_ratio = the
    largest |phi_hat|^2 at 0 < |k_ref| < 0.95 over the largest |phi_hat|^2
    (mean removed); defect_flag = either above its limit."""
    phi = np.asarray(state, dtype=np.float64)
    m = phi.mean()
    top = phi > m
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy or dx:
                top &= phi >= np.roll(phi, (dy, dx), axis=(0, 1))
    h = phi[top] - m
    spread = float((h.max() - h.min()) / h.mean()) if h.size else float('nan')
    power = np.abs(np.fft.fft2(phi - m)) ** 2
    kr = np.sqrt(k2_ref(cell))
    sub = (kr > 1e-9) & (kr < 0.95)
    ratio = float(power[sub].max() / power.max()) if power.max() > 0 else float('nan')
    return dict(peak_spread=spread, subq_ratio=ratio, n_peaks=int(top.sum()),
                defect_flag=bool(spread > PEAK_SPREAD_MAX or ratio > SUBQ_RATIO_MAX))
