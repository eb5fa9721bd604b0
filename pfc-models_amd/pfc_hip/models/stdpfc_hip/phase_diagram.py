"""Phase-diagram adapter for the hip stdPFC simulation (the driver's plugin API).

The reference's adapter names and meanings (models/stdpfc_cpu/phase_diagram.py:
SLUG, STRUCTURES :55-92, validate_* :101-143, build :220-256, get_state ..
free_energy :258-283, diagnose :286-296, liquid_* and the bounds :301-381) over
`BuildHipGPU_Simulation`.  stdPFC conventions: the field n may take either
sign; the ansatz amplitude is sign * scale * A_i in absolute units (the
triangle and the honeycomb share one scale, so at g = 0 the triangle at -nbar
mirrors the honeycomb at +nbar); fitness amplitudes are absolute.  `run(sim,
n)` advances n steps in one device call.
"""
from __future__ import annotations

import numpy as np

from ... import phase_diagram as _pd
from .. import _adapter as _A
from . import simulation as _simulation

SLUG = 'stdpfc_hip'

STRUCTURES = {
    'triangle': _A.Structure('triangle', _A.HEX_SETS, +1, 1.0 / 3, 2),
    'honeycomb': _A.Structure('honeycomb', _A.HEX_SETS, -1, 1.0 / 3, 4),
    'stripe': _A.Structure('stripe', _A.STRIPE_SETS, +1, 1.0 / 2, None),
}


def validate_params(params):
    return _A.validate_params(SLUG, params)


def validate_cell(cell):
    return _A.validate_cell(SLUG, cell)


def validate_ansatz(ansatz):
    return _A.validate_ansatz(SLUG, ansatz)


def _structure(structure):
    return STRUCTURES[structure] if isinstance(structure, str) else structure


def initial_field(structure, cell, ansatz, meanDensity):
    s = _structure(structure)
    k = s.amplitudeSign * s.amplitude
    return _A.three_mode(s, cell, meanDensity, k * ansatz['A1'], k * ansatz['A2'], k * ansatz['A3'])


def build(params, structure, cell, ansatz, temp, meanDensity, q00, dt, state=None, nThreads=1, precision=64):
    """Model + box at q00 + (the ansatz if state is None, else state) -> Simulation."""
    p = validate_params(params)
    cell = validate_cell(cell)
    ansatz = validate_ansatz(ansatz)
    model = _simulation.BuildStdPFCModelParameters(Gamma=p['Gamma'], beta=p['beta'], q0=p['q0'], epsilon=temp,
                                                    g=p['g'], dt=dt, meanDensity=meanDensity)
    sim = _simulation.BuildHipGPU_Simulation(model, _A.box(cell, q00), precision=precision)
    sim.State.n[...] = initial_field(structure, cell, ansatz, meanDensity) if state is None else state
    return sim


def get_state(sim):
    return np.array(sim.n.get(), dtype=np.float64)


def set_state(sim, state):
    sim.State.n[...] = state


def get_time(sim):
    return sim.t


def set_time(sim, t):
    sim.State.t = t


def get_dt(sim):
    return sim.Model.dt


def step(sim):
    sim.TimeStep()


def run(sim, n):
    sim.run(n)


def free_energy(sim):
    return sim.CalculateFreeEnergy()


def diagnose(sim, structure, cell):
    """fitness = 1 - mean((n - n_3mode)^2); amplitudes (A, B, C) = mean
    |n_hat| over the 1-, 2-, 3-mode sets per site; details:
    maxima_per_cell.  Fitness and amplitudes come from the device's spectral
    diagnostics (`SpectralFit`); `diagnose_state` computes them from a host
    field with host FFTs (the same values to rounding)."""
    fit = _fit(structure, cell)
    (fitness,), (amps,) = fit.assemble([sim._plan.diagnose(fit.ky, fit.kx)])
    return _pd.Diagnostics(fitness=float(fitness), amplitudes=tuple(float(a) for a in amps),
                           details={'maxima_per_cell': _A.maxima_per_cell(get_state(sim), validate_cell(cell))})


def _fit(structure, cell):
    """The SpectralFit of a structure on a cell (the device form of `diagnose_state`)."""
    return _A.spectral_fit(_structure(structure), validate_cell(cell), relative=False, nonnegative=False)


# -- batched evaluations (pfc_hip.phase_diagram.batched) ---------------------------
def open_batch(params, cell, capacity, precision=64):
    """A device batch of `capacity` evaluation slots on the cell's grid."""
    p = validate_params(params)
    model = _simulation.BuildStdPFCModelParameters(Gamma=p['Gamma'], beta=p['beta'], q0=p['q0'], epsilon=0.0,
                                                   g=p['g'], dt=1.0)
    return _A.SlotBatch('std', model, validate_cell(cell), capacity, _fit, precision=precision)


def ansatz_state(structure, cell, ansatz, meanDensity):
    """The fresh ansatz `build` starts from, as a host field."""
    return initial_field(structure, validate_cell(cell), validate_ansatz(ansatz), meanDensity)


def diagnose_state(state, structure, cell):
    """Fitness and amplitudes of a host field, with host FFTs."""
    return _A.diagnostics(np.asarray(state, dtype=np.float64), _structure(structure), validate_cell(cell))


# -- frozen field (pfc_hip.phase_diagram.frozen_q) ---------------------------------
def frozen_field(state, params, temp, cell):
    """f(q) and q_opt of `state` frozen on the grid while the box scale
    changes: the local part (the box mean, epsilon = temp, no gradient) plus
    beta/2 (q0^2 - q^2 k_ref^2)^2 |field_hat|^2 summed."""
    p = validate_params(params)
    n = np.asarray(state, dtype=np.float64)
    local = np.mean(temp / 2 * n ** 2 + p['g'] / 3 * n ** 3 + p['v0'] / 4 * n ** 4)
    return _A.FrozenField(state, p['beta'], p['q0'], local, validate_cell(cell))


# -- uniform state and bounds (closed forms) --------------------------------------
def liquid_free_energy(n, params, temp):
    """(beta q0^4 + eps)/2 n^2 + g/3 n^3 + v0/4 n^4."""
    p = validate_params(params)
    return 0.5 * (p['beta'] * p['q0'] ** 4 + temp) * n ** 2 + p['g'] / 3.0 * n ** 3 + p['v0'] / 4.0 * n ** 4


def liquid_chemical_potential(n, params, temp):
    """(beta q0^4 + eps) n + g n^2 + v0 n^3."""
    p = validate_params(params)
    return (p['beta'] * p['q0'] ** 4 + temp) * n + p['g'] * n ** 2 + p['v0'] * n ** 3


def _band(g, v0, eps_eff, width):
    disc = g ** 2 - 3 * v0 * eps_eff
    if not disc > 0:
        return None
    c = -g / (3 * v0)
    half = width * np.sqrt(disc) / v0
    return (c - half, c + half)


def critical_point(params):
    """(n_c, eps_c) = (-g/(3 v0), g^2/(3 v0))."""
    p = validate_params(params)
    return -p['g'] / (3 * p['v0']), p['g'] ** 2 / (3 * p['v0'])


def spinodal(params, temp):
    """k = q0 instability band eps + 2 g n + 3 v0 n^2 < 0, or None above eps_c."""
    p = validate_params(params)
    return _band(p['g'], p['v0'], temp, 1.0 / 3)


def liquid_spinodal(params, temp):
    """k = 0 concave band of the uniform branch, or None."""
    p = validate_params(params)
    return _band(p['g'], p['v0'], p['beta'] * p['q0'] ** 4 + temp, 1.0 / 3)


def existence_bounds(params, temp):
    """One-mode existence intervals per structure (saddle-node roots
    n_c -/+ sqrt(5 (g^2 - 3 v0 eps))/(6 v0) and the spinodal), or None."""
    p = validate_params(params)
    band = spinodal(p, temp)
    if band is None:
        return None
    sn = _band(p['g'], p['v0'], temp, np.sqrt(5.0) / 6)
    lo, hi = band
    return {'triangle': (sn[0], hi), 'stripe': (lo, hi), 'honeycomb': (lo, sn[1])}


def scan_window(structure, params, temp, pad=0.02):
    """The structure's existence interval widened by `pad`, or None."""
    b = existence_bounds(params, temp)
    if b is None:
        return None
    lo, hi = b[_structure(structure).name]
    return (lo - pad, hi + pad)
