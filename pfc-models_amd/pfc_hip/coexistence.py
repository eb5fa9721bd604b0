"""Coexistence of two free-energy branches of one conserved density.

A branch is the free energy per site f(phi) of one phase as a function of the
mean density phi: analytic (a uniform state's closed form) or tabulated (the
relaxed f of instances at a list of densities, as a sweep produces them).
Coexistence is the common tangent: densities phi_a, phi_b with equal chemical
potential mu = f'(phi) and equal grand potential omega = f - mu phi (the
Maxwell construction of phase_diagram/common_tangent.py:1-40 in the
reference; its `tabulated_branch` uses scipy's not-a-knot cubic spline,
:336-403, as this module does, and its `common_tangent`, :522-600, solves the
same two conditions).

Method.  A branch is split into convex pieces, the density intervals where
f'' > 0, on which mu(phi) is increasing and can be inverted.  For a pair of
pieces the difference omega_a(mu) - omega_b(mu) is sampled over the overlap
of their mu ranges and every sign change is refined with brentq in mu; the
inversions phi(mu) use brentq too.  Among several tangents the one with the
lowest omega (the supporting line under both branches) is returned.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.optimize import brentq


@dataclass
class Branch:
    name: str
    f: object          # callable phi -> f
    mu: object         # callable phi -> f'
    d2f: object        # callable phi -> f''
    lo: float
    hi: float
    pieces: tuple      # ((lo, hi), ...) where f'' > 0

    @property
    def nonconvex(self):
        out, x = [], self.lo
        for a, b in self.pieces:
            if a > x:
                out.append((x, a))
            x = b
        if x < self.hi:
            out.append((x, self.hi))
        return tuple(out)


def _convex_pieces(d2f, lo, hi, n=2001):
    x = np.linspace(lo, hi, n)
    d = np.asarray(d2f(x), dtype=float)
    pieces, start = [], None
    for i in range(n):
        pos = d[i] > 0
        if pos and start is None:
            start = x[i] if i == 0 else brentq(d2f, x[i - 1], x[i], xtol=1e-15)
        if not pos and start is not None:
            end = brentq(d2f, x[i - 1], x[i], xtol=1e-15) if d[i] < 0 else x[i]
            if end > start:
                pieces.append((start, end))
            start = None
    if start is not None and hi > start:
        pieces.append((start, hi))
    return tuple(pieces)


def analytic_branch(name, f, mu, lo, hi, d2f=None):
    """Closed-form branch on [lo, hi]; without d2f, f'' is a central
    difference of mu."""
    if d2f is None:
        h = 1e-6 * max(1.0, abs(hi - lo))

        def d2f(p):
            return (np.asarray(mu(np.asarray(p) + h)) - np.asarray(mu(np.asarray(p) - h))) / (2 * h)
    return Branch(name, f, mu, d2f, float(lo), float(hi), _convex_pieces(d2f, lo, hi))


def tabulated_branch(name, phi, f):
    """Branch from (phi, f) samples: a not-a-knot cubic spline of f, its
    derivatives giving mu and f''.  The spline's f'' is linear on each
    interval, so convex pieces are found exactly per interval."""
    phi = np.asarray(phi, dtype=float)
    f = np.asarray(f, dtype=float)
    order = np.argsort(phi)
    phi, f = phi[order], f[order]
    sp = CubicSpline(phi, f)
    d1, d2 = sp.derivative(1), sp.derivative(2)
    parts = []
    for i in range(len(phi) - 1):
        a, b = phi[i], phi[i + 1]
        ya, yb = float(d2(a)), float(d2(b))
        if ya > 0 and yb > 0:
            parts.append((a, b))
        elif ya > 0 or yb > 0:  # one sign change of the linear f''
            z = a + (b - a) * ya / (ya - yb)
            parts.append((a, z) if ya > 0 else (z, b))
    pieces = []
    for a, b in parts:
        if pieces and pieces[-1][1] == a:
            pieces[-1] = (pieces[-1][0], b)
        elif b > a:
            pieces.append((a, b))
    return Branch(name, sp, d1, d2, float(phi[0]), float(phi[-1]), tuple(pieces))


@dataclass
class Tangent:
    status: str                 # "ok", "no_overlap", "no_root"
    phi_a: float = float("nan")
    phi_b: float = float("nan")
    mu: float = float("nan")
    omega: float = float("nan")


def _phi_of_mu(br, piece, m):
    lo, hi = piece
    return brentq(lambda p: float(br.mu(p)) - m, lo, hi, xtol=1e-15, rtol=4 * np.finfo(float).eps)


def _omega(br, piece, m):
    p = _phi_of_mu(br, piece, m)
    return float(br.f(p)) - m * p, p


def common_tangent(a: Branch, b: Branch, samples=64) -> Tangent:
    """The common tangent of two branches (the lowest supporting one when
    several exist)."""
    found = []
    overlap_any = False
    for pa in a.pieces:
        ma = (float(a.mu(pa[0])), float(a.mu(pa[1])))
        for pb in b.pieces:
            mb = (float(b.mu(pb[0])), float(b.mu(pb[1])))
            lo, hi = max(ma[0], mb[0]), min(ma[1], mb[1])
            if not hi > lo:
                continue
            overlap_any = True
            # stay strictly inside the pieces' mu ranges
            eps = 1e-12 * max(1.0, abs(hi - lo))
            grid = np.linspace(lo + eps, hi - eps, samples)

            def g(m):
                return _omega(a, pa, m)[0] - _omega(b, pb, m)[0]

            vals = np.array([g(m) for m in grid])
            for i in range(samples - 1):
                if vals[i] == 0.0 or (vals[i] > 0) != (vals[i + 1] > 0):
                    m = grid[i] if vals[i] == 0.0 else brentq(g, grid[i], grid[i + 1], xtol=1e-16,
                                                             rtol=4 * np.finfo(float).eps)
                    om, xa = _omega(a, pa, m)
                    xb = _phi_of_mu(b, pb, m)
                    if abs(xa - xb) > 1e-9:
                        found.append(Tangent("ok", xa, xb, m, om))
    if found:
        return min(found, key=lambda t: t.omega)
    return Tangent("no_root" if overlap_any else "no_overlap")


# ---------------------------------------------------------------------------
# the stdPFC sweep's coexistence (liquid against the relaxed crystal branch)
# ---------------------------------------------------------------------------
def std_liquid(r, lo=-1.0, hi=1.0):
    """Uniform stdPFC state, F = psi/2 (r + (1 + lap)^2) psi + psi^4/4 at k = 0:
    f = (r + 1)/2 psi^2 + psi^4/4."""
    return analytic_branch("liquid", lambda p: 0.5 * (r + 1) * np.asarray(p) ** 2 + 0.25 * np.asarray(p) ** 4,
                           lambda p: (r + 1) * np.asarray(p) + np.asarray(p) ** 3, lo, hi,
                           d2f=lambda p: (r + 1) + 3 * np.asarray(p) ** 2)


def std_sweep_coexistence(r, psibar, fe, crystal_margin=1e-6, min_points=4):
    """Liquid-crystal coexistence at one r from relaxed (psibar, fe) pairs:
    instances at least `crystal_margin` below the liquid's f form the crystal
    branch.  Returns a Tangent (status "too_few" without enough crystal
    points)."""
    psibar = np.asarray(psibar, dtype=float)
    fe = np.asarray(fe, dtype=float)
    liq = std_liquid(r)
    crystal = fe < np.asarray(liq.f(psibar)) - crystal_margin
    if crystal.sum() < min_points:
        return Tangent("too_few")
    tri = tabulated_branch("crystal", psibar[crystal], fe[crystal])
    lo = float(min(psibar.min(), -1.0))
    hi = float(max(psibar.max(), 1.0))
    return common_tangent(std_liquid(r, lo, hi), tri)
