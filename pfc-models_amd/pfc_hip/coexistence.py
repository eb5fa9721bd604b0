"""Liquid-crystal coexistence of the stdPFC sweep.

The common-tangent construction itself (branches, pairwise tangents,
envelopes, Maxwell lines) is `pfc_hip.phase_diagram.common_tangent`, the
restatement of the reference's phase_diagram/common_tangent.py; this module
adds the uniform stdPFC state as an analytic branch and the sweep's use of it:
at one r, the relaxed (psibar, f) pairs that lie below the liquid form the
crystal branch, whose tangent with the liquid gives the coexisting densities.
"""
from __future__ import annotations

import numpy as np

from .phase_diagram.common_tangent import (  # noqa: F401  (re-exported)
    Branch, TangentResult, analytic_branch, common_tangent, lower_envelope, maxwell_lines, tabulated_branch)


def std_liquid(r, lo=-1.0, hi=1.0):
    """Uniform stdPFC state, F = psi/2 (r + (1 + lap)^2) psi + psi^4/4 at k = 0:
    f = (r + 1)/2 psi^2 + psi^4/4."""
    return analytic_branch("liquid", lambda p: 0.5 * (r + 1) * np.asarray(p) ** 2 + 0.25 * np.asarray(p) ** 4,
                           lambda p: (r + 1) * np.asarray(p) + np.asarray(p) ** 3, lo, hi,
                           d2f=lambda p: (r + 1) + 3 * np.asarray(p) ** 2)


def std_sweep_coexistence(r, psibar, fe, crystal_margin=1e-6, min_points=4):
    """Liquid-crystal coexistence at one r from relaxed (psibar, fe) pairs:
    instances at least `crystal_margin` below the liquid's f form the crystal
    branch.  Returns a TangentResult (status "too_few" without `min_points`
    crystal points)."""
    psibar = np.asarray(psibar, dtype=float)
    fe = np.asarray(fe, dtype=float)
    liq = std_liquid(r)
    crystal = fe < np.asarray(liq.f(psibar)) - crystal_margin
    if crystal.sum() < min_points:
        return TangentResult("liquid", "crystal", "too_few",
                             message=f"{int(crystal.sum())} relaxed states below the liquid (< {min_points})")
    tri = tabulated_branch("crystal", psibar[crystal], fe[crystal], min_points=min_points)
    lo = float(min(psibar.min(), -1.0))
    hi = float(max(psibar.max(), 1.0))
    return common_tangent(std_liquid(r, lo, hi), tri)


def liquid_crystal_pair_ok(t, min_gap=1e-3):
    """Whether a liquid-crystal tangent of the stdPFC sweep is a coexistence:
    at psibar < 0 the uniform liquid is the lower density, so a pair with
    phi_liquid >= phi_crystal - min_gap (reversed, or a degenerate sliver) is
    spurious: it comes from crystal free energies that are not converged
    relaxations (noise-seeded states still nucleating), not from the model's
    phase diagram."""
    return bool(t.ok and t.phi_b - t.phi_a >= min_gap)
