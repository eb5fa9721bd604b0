"""Phase-diagram adapter for the hip logPFC simulation (the driver's plugin API).

The reference's adapter names and meanings (models/logpfc_cpu/phase_diagram.py:
SLUG, STRUCTURES :68-101, validate_* :112-155, build :238-278, get_state ..
free_energy :281-306, diagnose :309-321, liquid_* and the bounds :324-462)
over `BuildHipGPU_Simulation` in FP64 with the unclamped log, as the
reference's logPFC step uses.  logPFC conventions: the density phi is
positive; the ansatz amplitude is sign * amplitudeMax * phibar * A_i, and an
ansatz with a negative minimum is rescaled about phibar to a zero minimum;
fitness amplitudes are in units of phibar.  `run(sim, n)` advances n steps in
one device call.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import brentq

from ... import phase_diagram as _pd
from .. import _adapter as _A
from . import simulation as _simulation

SLUG = 'logpfc_hip'

STRUCTURES = {
    'triangle': _A.Structure('triangle', _A.HEX_SETS, +1, 1.0 / 3, 2),
    'honeycomb': _A.Structure('honeycomb', _A.HEX_SETS, -1, 1.0 / 6, 4),
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
    k = s.amplitudeSign * s.amplitude * meanDensity
    return _A.three_mode(s, cell, meanDensity, k * ansatz['A1'], k * ansatz['A2'], k * ansatz['A3'],
                         nonnegative=True)


def build(params, structure, cell, ansatz, temp, meanDensity, q00, dt, state=None, nThreads=1, precision=64):
    """Model + box at q00 + (the ansatz if state is None, else state) -> Simulation."""
    p = validate_params(params)
    cell = validate_cell(cell)
    ansatz = validate_ansatz(ansatz)
    model = _simulation.BuildLogPFCModelParameters(Gamma=p['Gamma'], beta=p['beta'], q0=p['q0'], epsilon=temp,
                                                    g=p['g'], dt=dt, meanDensity=meanDensity)
    sim = _simulation.BuildHipGPU_Simulation(model, _A.box(cell, q00), precision=precision)
    sim.State.phi[...] = initial_field(structure, cell, ansatz, meanDensity) if state is None else state
    return sim


def get_state(sim):
    return np.array(sim.phi.get(), dtype=np.float64)


def set_state(sim, state):
    sim.State.phi[...] = state


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
    """fitness = 1 - mean((phi - phi_3mode)^2); amplitudes (A, B, C) = mean
    |phi_hat| over the 1-, 2-, 3-mode sets per site in units of phibar; details:
    maxima_per_cell.  Fitness and amplitudes come from the device's spectral
    diagnostics (`SpectralFit`); `diagnose_state` computes them from a host
    field with host FFTs (the same values to rounding)."""
    fit = _fit(structure, cell)
    (fitness,), (amps,) = fit.assemble([sim._plan.diagnose(fit.ky, fit.kx)])
    return _pd.Diagnostics(fitness=float(fitness), amplitudes=tuple(float(a) for a in amps),
                           details={'maxima_per_cell': _A.maxima_per_cell(get_state(sim), validate_cell(cell))})


def _fit(structure, cell):
    """The SpectralFit of a structure on a cell (the device form of `diagnose_state`)."""
    return _A.spectral_fit(_structure(structure), validate_cell(cell), relative=True, nonnegative=True)


# -- batched evaluations (pfc_hip.phase_diagram.batched) ---------------------------
def open_batch(params, cell, capacity, precision=64):
    """A device batch of `capacity` evaluation slots on the cell's grid."""
    p = validate_params(params)
    model = _simulation.BuildLogPFCModelParameters(Gamma=p['Gamma'], beta=p['beta'], q0=p['q0'], epsilon=0.0,
                                                   g=p['g'], dt=1.0)
    return _A.SlotBatch('log', model, validate_cell(cell), capacity, _fit, precision=precision)


def ansatz_state(structure, cell, ansatz, meanDensity):
    """The fresh ansatz `build` starts from, as a host field."""
    return initial_field(structure, validate_cell(cell), validate_ansatz(ansatz), meanDensity)


def diagnose_state(state, structure, cell):
    """Fitness and amplitudes of a host field, with host FFTs."""
    return _A.diagnostics(np.asarray(state, dtype=np.float64), _structure(structure), validate_cell(cell),
                          relative=True, nonnegative=True)


# -- frozen field (pfc_hip.phase_diagram.frozen_q) ---------------------------------
def frozen_field(state, params, temp, cell):
    """f(q) and q_opt of `state` frozen on the grid while the box scale
    changes: the local part (the box mean, epsilon = temp, no gradient) plus
    beta/2 (q0^2 - q^2 k_ref^2)^2 |field_hat|^2 summed."""
    p = validate_params(params)
    phi = np.asarray(state, dtype=np.float64)
    local = np.mean(phi * np.log(phi) + temp / 2 * phi ** 2 + p['g'] / 3 * phi ** 3 + p['v0'] / 4 * phi ** 4)
    return _A.FrozenField(state, p['beta'], p['q0'], local, validate_cell(cell))


def defect_diagnostics(state, cell):
    """peak_spread, subq_ratio, n_peaks, defect_flag of a relaxed field (record only)."""
    return _A.defect_diagnostics(state, validate_cell(cell))


# -- uniform state and bounds ---------------------------------------------------------
def liquid_free_energy(phi, params, temp):
    """phi ln phi + (beta q0^4 + eps)/2 phi^2 + g/3 phi^3 + v0/4 phi^4."""
    p = validate_params(params)
    return (phi * np.log(phi) + 0.5 * (p['beta'] * p['q0'] ** 4 + temp) * phi ** 2 + p['g'] / 3.0 * phi ** 3
            + p['v0'] / 4.0 * phi ** 4)


def liquid_chemical_potential(phi, params, temp):
    """df/dphi of the uniform state, the +1 included."""
    p = validate_params(params)
    return 1 + np.log(phi) + (p['beta'] * p['q0'] ** 4 + temp) * phi + p['g'] * phi ** 2 + p['v0'] * phi ** 3


def _h(x, g, v0):  # local f'' without eps
    return 1 / x + 2 * g * x + 3 * v0 * x ** 2


def _f3(x, g, v0):
    return -1 / x ** 2 + 2 * g + 6 * v0 * x


def _f4(x, g, v0):
    return 2 / x ** 3 + 6 * v0


def _phi_c(g, v0):
    """The minimiser of h on phi > 0 (root of the increasing f3)."""
    hi = 1.0
    while _f3(hi, g, v0) <= 0:
        hi *= 2
    return brentq(lambda x: _f3(x, g, v0), 1e-8, hi, xtol=1e-15, maxiter=500)


def _roots_around(F, c):
    """The two roots of a convex F either side of its minimiser c, or None if F(c) >= 0."""
    if not F(c) < 0:
        return None
    lo, hi = c, c
    while F(lo) < 0:
        lo /= 2
    while F(hi) < 0:
        hi *= 2
    return brentq(F, lo, c, xtol=1e-15, maxiter=500), brentq(F, c, hi, xtol=1e-15, maxiter=500)


def critical_point(params):
    """(phi_c, eps_c) with eps_c = -h(phi_c)."""
    p = validate_params(params)
    c = _phi_c(p['g'], p['v0'])
    return c, -_h(c, p['g'], p['v0'])


def spinodal(params, temp):
    """k = q0 instability band: eps + h(phi) < 0, or None at and above eps_c."""
    p = validate_params(params)
    return _roots_around(lambda x: temp + _h(x, p['g'], p['v0']), _phi_c(p['g'], p['v0']))


def liquid_spinodal(params, temp):
    """k = 0 band where the uniform branch is concave, or None."""
    p = validate_params(params)
    return _roots_around(lambda x: p['beta'] * p['q0'] ** 4 + temp + _h(x, p['g'], p['v0']),
                         _phi_c(p['g'], p['v0']))


def _first_sign_change(G, start, stop, n=4000):
    grid = np.geomspace(start, stop, n)
    vals = np.array([G(x) for x in grid])
    k = np.flatnonzero(vals <= 0)
    if k.size == 0:
        return None
    return brentq(G, grid[k[0] - 1], grid[k[0]], xtol=1e-15, maxiter=500)


def existence_bounds(params, temp):
    """One-mode existence intervals per structure from the saddle-node line
    eps_SN = -h + f3^2/(10 f4) and the spinodal, or None at and above eps_c."""
    p = validate_params(params)
    g, v0 = p['g'], p['v0']
    band = spinodal(p, temp)
    if band is None:
        return None
    c = _phi_c(g, v0)

    def G(x):
        return -_h(x, g, v0) + 0.1 * _f3(x, g, v0) ** 2 / _f4(x, g, v0) - temp
    lo, hi = band
    return {'triangle': (_first_sign_change(G, c, c * 1e-4), hi), 'stripe': (lo, hi),
            'honeycomb': (lo, _first_sign_change(G, c, c * 1e4))}


def scan_window(structure, params, temp, pad=0.02):
    """The existence interval widened by `pad` (the lower end not below half of it), or None."""
    b = existence_bounds(params, temp)
    if b is None:
        return None
    lo, hi = b[_structure(structure).name]
    return (max(lo - pad, 0.5 * lo), hi + pad)
