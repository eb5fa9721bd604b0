"""Coexistence between free-energy branches of one conserved density.

Host-side consumer of relaxed free energies (q0* scans, sweeps); restates
`phase_diagram/common_tangent.py` of the reference (API at :98-113, branches
:127-403, pairwise tangent :410-600, lower envelope :603-845, Maxwell lines
:847-989) so that tables, envelopes and Maxwell-line CSVs produced from this
framework's fmin curves match the reference's.

A branch is the free energy per site f(phi) of one phase against the mean
density phi: analytic (a uniform state's closed form) or tabulated (fmin at a
list of densities, NaN where a relaxation failed).  A tabulated branch is one
not-a-knot cubic spline per run of finite points ("segment"); the NaN runs
between segments are its gaps.  Every segment is split at the zeros of f''
into convex pieces (f'' > 0: mu = f' strictly increasing, so phi(mu) is single
valued) and nonconvex intervals.

Common tangent of pieces A and B: a chemical potential mu with equal grand
potential omega = f - mu phi on both, i.e. a root of
omega_A(mu) - omega_B(mu) on the overlap of their mu ranges.  The difference
is scanned at `n_scan` points and every sign change refined by brentq in mu
(xtol 1e-15, rtol 4 eps), each omega evaluation inverting mu(phi) on its piece
by brentq.  Statuses: ok, no_overlap (mu ranges disjoint), no_root (convex mu
ranges overlap, no sign change; also a self-pair of a branch convex
throughout), non_convex (no convex piece, or overlap only through concave
parts).  Among several tangents the primary one is a supporting one (no branch
of the pair below its line by more than atol), then the one nearest
`mu_guess`, else the lowest mu.

Lower envelope at one temperature: the lower convex hull of every branch
sampled on a density grid gives the single-phase stretches in density order;
neighbouring stretches are bridged by the ordered tangent of their pieces;
then, until nothing changes, a window whose two bridges cross is removed and a
branch dipping below a bridge's line is inserted as a window.  Every other
tangent is metastable, labelled with a branch below its line.  Maxwell lines:
envelopes over temperatures as row tables (CSV), with triple points bracketed
where a window between L and R at one temperature is gone (L and R adjacent)
at the next, its signed width interpolated to zero from the metastable L-X and
X-R tangents when both exist.
"""
from __future__ import annotations

import csv
import math
from collections.abc import Callable, Iterable, Mapping, Sequence
from dataclasses import asdict, dataclass, field, replace

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.optimize import brentq

__all__ = [
    "STATUSES", "Segment", "Piece", "Branch", "TangentResult", "Window", "Envelope", "TripleBracket",
    "MaxwellLines", "analytic_branch", "tabulated_branch", "common_tangent", "lower_envelope", "maxwell_lines",
]

STATUSES = ("ok", "no_overlap", "no_root", "non_convex")

_RTOL = 4.0 * np.finfo(float).eps
_XTOL = 1e-15


def _val(fn, x) -> float:
    """fn at one density, as a Python float (fn takes numpy arrays)."""
    return float(np.asarray(fn(np.array([float(x)])), dtype=float).ravel()[0])


# ----------------------------------------------------------------------------- branches
@dataclass(frozen=True)
class Segment:
    """A contiguous density interval of a branch with its own f, mu (= f'),
    f'' (or None) and q0* (or None); `nodes` are a tabulated segment's table
    densities."""
    lo: float
    hi: float
    f: Callable
    mu: Callable
    d2f: Callable | None = None
    q0star: Callable | None = None
    nodes: np.ndarray | None = field(default=None, repr=False, compare=False)


@dataclass(frozen=True)
class Piece:
    """Convex part (f'' > 0) of segment `segment`, with mu at its ends."""
    segment: int
    lo: float
    hi: float
    mu_lo: float
    mu_hi: float


@dataclass(frozen=True)
class Branch:
    """Free energy of one phase over one or more density segments.

    `pieces` (convex parts, density order) are what tangents use; `nonconvex`
    lists the (lo, hi) intervals with f'' <= 0; `gaps` the density gaps between
    segments (NaN runs of a table); `dropped` the densities of finite runs too
    short to spline."""
    name: str
    kind: str
    segments: tuple
    pieces: tuple
    nonconvex: tuple
    gaps: tuple = ()
    dropped: tuple = ()
    _samples: tuple = field(default=(), repr=False, compare=False)

    @property
    def lo(self) -> float:
        return self.segments[0].lo

    @property
    def hi(self) -> float:
        return self.segments[-1].hi

    @property
    def has_q0star(self) -> bool:
        return any(s.q0star is not None for s in self.segments)

    @property
    def mu_range(self) -> tuple:
        """(min, max) of mu over the dense samples, concave parts included."""
        m = self._samples[2]
        return float(np.min(m)), float(np.max(m))

    def _where(self, which, phi):
        x = np.asarray(phi, dtype=float)
        flat = np.atleast_1d(x).ravel()
        out = np.full(flat.shape, np.nan)
        for seg in self.segments:
            fn = getattr(seg, which)
            if fn is None:
                continue
            inside = (flat >= seg.lo) & (flat <= seg.hi)
            if inside.any():
                out[inside] = np.asarray(fn(flat[inside]), dtype=float)
        return float(out[0]) if x.ndim == 0 else out.reshape(x.shape)

    def f(self, phi):
        """f(phi); NaN outside the segments."""
        return self._where("f", phi)

    def mu(self, phi):
        """mu(phi) = f'(phi); NaN outside the segments."""
        return self._where("mu", phi)

    def q0star(self, phi):
        """q0*(phi) where given; NaN elsewhere."""
        return self._where("q0star", phi)

    def piece_at(self, phi: float):
        """Index of the convex piece holding `phi`, or None."""
        return next((k for k, p in enumerate(self.pieces) if p.lo <= phi <= p.hi), None)


def _dense_samples(segments, n):
    """(phi, f, mu, segment) on n uniform points per segment plus its table nodes."""
    cols = ([], [], [], [])
    for k, seg in enumerate(segments):
        x = np.linspace(seg.lo, seg.hi, n)
        if seg.nodes is not None:
            x = np.union1d(x, seg.nodes)
        cols[0].append(x)
        cols[1].append(np.asarray(seg.f(x), dtype=float))
        cols[2].append(np.asarray(seg.mu(x), dtype=float))
        cols[3].append(np.full(x.shape, k))
    return tuple(np.concatenate(c) for c in cols)


def _split(segments, cuts, is_convex):
    """Cut every segment at its candidate f'' zeros and label each interval
    convex (a Piece; touching convex intervals merge) or nonconvex."""
    pieces, concave = [], []
    for k, (seg, cut) in enumerate(zip(segments, cuts)):
        edges = np.union1d(np.unique(np.clip(np.asarray(cut, dtype=float), seg.lo, seg.hi)), [seg.lo, seg.hi])
        tiny = 1e-12 * max(seg.hi - seg.lo, 1e-300)
        for u, v in zip(edges[:-1], edges[1:]):
            u, v = float(u), float(v)
            if v - u <= tiny:
                continue
            if is_convex(seg, u, v):
                mu_u, mu_v = _val(seg.mu, u), _val(seg.mu, v)
                if pieces and pieces[-1].segment == k and pieces[-1].hi == u:
                    last = pieces.pop()
                    pieces.append(Piece(k, last.lo, v, last.mu_lo, mu_v))
                else:
                    pieces.append(Piece(k, u, v, mu_u, mu_v))
            elif concave and concave[-1][1] == u:
                concave[-1] = (concave[-1][0], v)
            else:
                concave.append((u, v))
    return tuple(pieces), tuple(concave)


def analytic_branch(name: str, f: Callable, mu: Callable, lo: float, hi: float, *, d2f: Callable | None = None,
                    q0star: Callable | None = None, n_sample: int = 2049) -> Branch:
    """Branch of vectorised callables on [lo, hi] (mu = f' is the caller's to
    keep consistent).  Convex pieces come from the sign of d2f when given, its
    sign changes on n_sample points refined by brentq; otherwise from where mu
    stops increasing on those points, each turn refined as a root of a central
    difference of mu."""
    lo, hi = float(lo), float(hi)
    if not hi > lo:
        raise ValueError(f"analytic_branch {name!r}: need lo < hi, got {lo}, {hi}")
    seg = Segment(lo, hi, f, mu, d2f, q0star)
    x = np.linspace(lo, hi, n_sample)
    cuts = []
    if d2f is not None:
        s = np.sign(np.asarray(d2f(x), dtype=float))
        for i in np.flatnonzero(s[:-1] != s[1:]):
            try:
                cuts.append(brentq(lambda p: _val(d2f, p), x[i], x[i + 1], xtol=_XTOL, rtol=_RTOL))
            except ValueError:
                cuts.append(0.5 * (x[i] + x[i + 1]))

        def is_convex(sg, u, v):
            return _val(sg.d2f, 0.5 * (u + v)) > 0.0
    else:
        rising = np.sign(np.diff(np.asarray(mu(x), dtype=float)))
        h = 1e-6 * max(1.0, abs(lo), abs(hi))

        def slope(p):
            a, b = max(lo, p - h), min(hi, p + h)
            return (_val(mu, b) - _val(mu, a)) / (b - a)

        for i in np.flatnonzero(rising[:-1] != rising[1:]):
            try:
                cuts.append(brentq(slope, x[i], x[i + 2], xtol=1e-14, rtol=_RTOL))
            except ValueError:
                cuts.append(x[i + 1])

        def is_convex(sg, u, v):
            return _val(sg.mu, v) > _val(sg.mu, u)

    pieces, concave = _split((seg,), [cuts], is_convex)
    return Branch(name, "analytic", (seg,), pieces, concave, _samples=_dense_samples((seg,), n_sample))


def tabulated_branch(name: str, phi, f, q0star=None, *, min_points: int = 4, bc_type="not-a-knot",
                     n_sample: int = 1025) -> Branch:
    """Branch of a table f(phi) (NaN where not converged), e.g. fmin of a q0*
    scan.  Sorted by phi; every run of >= min_points consecutive finite points
    (phi, f and, when given, q0*) becomes a segment splined with `bc_type`;
    shorter runs are listed in `dropped`.  f'' of a cubic spline is piecewise
    linear, so the convex pieces are exact."""
    phi = np.asarray(phi, dtype=float).ravel()
    f = np.asarray(f, dtype=float).ravel()
    q = None if q0star is None else np.asarray(q0star, dtype=float).ravel()
    if phi.shape != f.shape or (q is not None and q.shape != phi.shape):
        raise ValueError(f"tabulated_branch {name!r}: phi, f, q0star must have equal length")
    order = np.argsort(phi, kind="stable")
    phi, f = phi[order], f[order]
    if q is not None:
        q = q[order]
    if np.any(np.diff(phi[np.isfinite(phi)]) == 0):
        raise ValueError(f"tabulated_branch {name!r}: repeated phi values")
    ok = np.isfinite(phi) & np.isfinite(f)
    if q is not None:
        ok &= np.isfinite(q)
    # maximal runs of usable points
    edges = np.flatnonzero(np.diff(np.concatenate(([0], ok.astype(np.int8), [0]))))
    segments, dropped, cuts = [], [], []
    for i, j in zip(edges[0::2], edges[1::2]):
        if j - i < min_points:
            dropped.extend(float(v) for v in phi[i:j])
            continue
        x = phi[i:j]
        spl = CubicSpline(x, f[i:j], bc_type=bc_type)
        qs = None if q is None else CubicSpline(x, q[i:j], bc_type=bc_type)
        d2 = spl.derivative(2)
        segments.append(Segment(float(x[0]), float(x[-1]), spl, spl.derivative(1), d2, qs, x))
        r = np.asarray(d2.roots(extrapolate=False), dtype=float)
        cuts.append(r[np.isfinite(r)])
    if not segments:
        raise ValueError(f"tabulated_branch {name!r}: no run of >= {min_points} finite points")

    def is_convex(sg, u, v):
        return float(sg.d2f(0.5 * (u + v))) > 0.0

    pieces, concave = _split(segments, cuts, is_convex)
    gaps = tuple((a.hi, b.lo) for a, b in zip(segments[:-1], segments[1:]))
    return Branch(name, "tabulated", tuple(segments), pieces, concave, gaps, tuple(dropped),
                  _samples=_dense_samples(segments, n_sample))


# ------------------------------------------------------------------------- pairwise tangent
@dataclass(frozen=True)
class TangentResult:
    """Common tangent between branches `a` and `b` (names), or the reason for none.

    phi_a, phi_b: coexisting densities on a and b in argument order.
    residual_mu_a/b = mu(phi) - mu on each side, residual_omega = omega_a -
    omega_b; piece_a/b index `Branch.pieces`.  `supporting`: neither branch
    dips below the line by more than atol.  `stable`/`below` are set by
    lower_envelope (bridge or metastable, and a branch under a metastable
    line).  `candidates`: every tangent found (primary included) when the
    search was not restricted to one piece pair."""
    a: str
    b: str
    status: str
    phi_a: float = math.nan
    phi_b: float = math.nan
    mu: float = math.nan
    omega: float = math.nan
    q0star_a: float | None = None
    q0star_b: float | None = None
    residual_mu_a: float = math.nan
    residual_mu_b: float = math.nan
    residual_omega: float = math.nan
    piece_a: int | None = None
    piece_b: int | None = None
    same_branch: bool = False
    across_gap: bool = False
    supporting: bool | None = None
    stable: bool | None = None
    below: str | None = None
    message: str = ""
    candidates: tuple = field(default=(), repr=False)

    @property
    def ok(self) -> bool:
        return self.status == "ok"

    @property
    def residual(self) -> float:
        """Largest absolute residual (NaN unless ok)."""
        return max(abs(self.residual_mu_a), abs(self.residual_mu_b), abs(self.residual_omega))

    def as_dict(self) -> dict:
        d = asdict(self)
        d.pop("candidates")
        return d


def _phi_at(br: Branch, k: int, mu: float) -> float:
    """phi with mu(phi) = mu on piece k (clamped to the piece's ends)."""
    p = br.pieces[k]
    if mu <= p.mu_lo:
        return p.lo
    if mu >= p.mu_hi:
        return p.hi
    m = br.segments[p.segment].mu
    return brentq(lambda x: _val(m, x) - mu, p.lo, p.hi, xtol=_XTOL, rtol=_RTOL)


def _grand(br: Branch, k: int, mu: float):
    """(omega, phi) on piece k at chemical potential mu."""
    x = _phi_at(br, k, mu)
    return _val(br.segments[br.pieces[k].segment].f, x) - mu * x, x


def _sign_changes(fn, lo, hi, n):
    """Roots of fn on [lo, hi]: exact zeros on an n-point grid and brentq in
    every bracketing grid interval, merged within 1e-12 relative."""
    grid = np.linspace(lo, hi, max(int(n), 2))
    vals = [fn(m) for m in grid]
    roots = []
    for k, m in enumerate(grid):
        if vals[k] == 0.0:
            roots.append(float(m))
        elif k + 1 < len(grid) and vals[k] * vals[k + 1] < 0.0:
            roots.append(brentq(fn, m, grid[k + 1], xtol=_XTOL, rtol=_RTOL))
    out = []
    for r in sorted(roots):
        if not out or abs(r - out[-1]) > 1e-12 * (1.0 + abs(r)):
            out.append(r)
    return out


def _below_line(br: Branch, mu: float, phi0: float, f0: float):
    """(min of f - line, phi there) over the branch's dense samples, for the
    line through (phi0, f0) with slope mu."""
    ph, ff = br._samples[0], br._samples[1]
    d = ff - (f0 + mu * (ph - phi0))
    k = int(np.argmin(d))
    return float(d[k]), float(ph[k])


def _tangent_at(a, i, b, j, mu, same, atol):
    sa, sb = a.segments[a.pieces[i].segment], b.segments[b.pieces[j].segment]
    om_a, xa = _grand(a, i, mu)
    om_b, xb = _grand(b, j, mu)
    fa = om_a + mu * xa
    dip = min(_below_line(a, mu, xa, fa)[0], _below_line(b, mu, xa, fa)[0])
    return TangentResult(
        a=a.name, b=b.name, status="ok", phi_a=float(xa), phi_b=float(xb), mu=float(mu), omega=float(om_a),
        q0star_a=None if sa.q0star is None else _val(sa.q0star, xa),
        q0star_b=None if sb.q0star is None else _val(sb.q0star, xb),
        residual_mu_a=_val(sa.mu, xa) - mu, residual_mu_b=_val(sb.mu, xb) - mu,
        residual_omega=float(om_a - om_b), piece_a=i, piece_b=j, same_branch=same,
        across_gap=bool(same and a.pieces[i].segment != b.pieces[j].segment),
        supporting=bool(dip >= -atol))


def common_tangent(a: Branch, b: Branch, *, pieces=None, ordered: bool = False, mu_guess: float | None = None,
                   n_scan: int = 33, atol: float = 1e-10) -> TangentResult:
    """Common tangent (equal mu and omega) of branches a and b over every pair
    of convex pieces (distinct pieces when `a is b`: a self-tangent across a
    concave interval).  `pieces=(i, j)` restricts the search to one pair,
    `ordered` keeps only phi_a < phi_b, `mu_guess` prefers the candidate
    nearest that mu.  Never raises for a pair without a tangent."""
    same = a is b
    for br in (a, b):
        if not br.pieces:
            return TangentResult(a.name, b.name, "non_convex", same_branch=same,
                                 message=f"branch {br.name!r} has no convex piece")
    if pieces is not None:
        pairs = [tuple(pieces)]
    elif same:
        pairs = [(i, j) for i in range(len(a.pieces)) for j in range(i + 1, len(a.pieces))]
    else:
        pairs = [(i, j) for i in range(len(a.pieces)) for j in range(len(b.pieces))]
    found, overlap = [], False
    for i, j in pairs:
        lo = max(a.pieces[i].mu_lo, b.pieces[j].mu_lo)
        hi = min(a.pieces[i].mu_hi, b.pieces[j].mu_hi)
        if not lo < hi:
            continue
        overlap = True

        def domega(m, i=i, j=j):
            return _grand(a, i, m)[0] - _grand(b, j, m)[0]

        found += [_tangent_at(a, i, b, j, m, same, atol) for m in _sign_changes(domega, lo, hi, n_scan)]
    if ordered:
        found = [c for c in found if c.phi_a < c.phi_b]
    if found:
        found.sort(key=lambda c: (not c.supporting, abs(c.mu - mu_guess) if mu_guess is not None else c.mu))
        best = replace(found[0], candidates=tuple(found) if pieces is None else ())
        if len(found) > 1:
            best = replace(best, message=f"{len(found)} tangents found; primary chosen by supporting, "
                                         "then mu_guess/lowest mu")
        return best
    if same and len(a.pieces) < 2:
        return TangentResult(a.name, b.name, "no_root", same_branch=True,
                             message="branch is convex on one piece; no self-tangent")
    (alo, ahi), (blo, bhi) = a.mu_range, b.mu_range
    if not max(alo, blo) < min(ahi, bhi):
        return TangentResult(a.name, b.name, "no_overlap", same_branch=same,
                             message=f"mu ranges [{alo:.6g}, {ahi:.6g}] and [{blo:.6g}, {bhi:.6g}] are disjoint")
    if overlap:
        return TangentResult(a.name, b.name, "no_root", same_branch=same,
                             message="no sign change of omega_a - omega_b on the convex mu overlap"
                                     + (" with phi_a < phi_b" if ordered else ""))
    return TangentResult(a.name, b.name, "non_convex", same_branch=same,
                         message="mu ranges overlap only through concave parts")


# --------------------------------------------------------------------------- lower envelope
@dataclass(frozen=True)
class Window:
    """Densities where phase `name` (branches[branch].pieces[piece]) alone is stable."""
    name: str
    branch: int
    piece: int
    phi_lo: float
    phi_hi: float

    @property
    def width(self) -> float:
        return self.phi_hi - self.phi_lo


@dataclass(frozen=True)
class Envelope:
    """Lower envelope at one temperature: windows[k] | coexistence[k] |
    windows[k+1] in density order; `metastable` holds every other tangent
    (stable=False, `below` a branch under its line); hull_phi/hull_f are the
    sampled hull vertices."""
    temp: float | None
    names: tuple
    windows: tuple
    coexistence: tuple
    metastable: tuple
    warnings: tuple
    hull_phi: np.ndarray = field(repr=False, compare=False, default=None)
    hull_f: np.ndarray = field(repr=False, compare=False, default=None)


def _hull_lower(x, y):
    """Indices of the lower convex hull of points sorted by x (monotone chain)."""
    h = []
    for k in range(len(x)):
        while len(h) >= 2:
            o, p = h[-2], h[-1]
            if (x[p] - x[o]) * (y[k] - y[o]) - (y[p] - y[o]) * (x[k] - x[o]) <= 0.0:
                h.pop()
            else:
                break
        h.append(k)
    return h


def _closest_piece(br: Branch, phi: float, seg: int) -> int:
    """Index of the piece of segment `seg` nearest `phi` (-1 if it has none)."""
    best, dist = -1, math.inf
    for k, p in enumerate(br.pieces):
        if p.segment != seg:
            continue
        d = 0.0 if p.lo <= phi <= p.hi else min(abs(phi - p.lo), abs(phi - p.hi))
        if d < dist:
            best, dist = k, d
    return best


def lower_envelope(branches: Sequence[Branch], phi_grid=None, *, temp: float | None = None, n: int = 4001,
                   n_scan: int = 33, atol: float = 1e-10, max_iter: int | None = None) -> Envelope:
    """Coexisting pairs and single-phase windows of all branches at one
    temperature (branch names unique).  The hull of the branches sampled on
    `phi_grid` (default: n uniform points over their union, plus segment ends)
    orders the stretches; bridges are ordered tangents of neighbouring pieces;
    crossed windows are removed and dipping branches inserted until stable."""
    branches = list(branches)
    if not branches:
        raise ValueError("lower_envelope: no branches")
    names = tuple(b.name for b in branches)
    if len(set(names)) != len(names):
        raise ValueError(f"lower_envelope: branch names must be unique, got {names}")
    warns = []
    grid = (np.linspace(min(b.lo for b in branches), max(b.hi for b in branches), n) if phi_grid is None
            else np.asarray(phi_grid, dtype=float))
    # every branch on the grid (plus its segment ends), labelled (branch, piece)
    ph, ff, lb, lp = [], [], [], []
    for bi, br in enumerate(branches):
        for si, seg in enumerate(br.segments):
            x = np.union1d(grid[(grid >= seg.lo) & (grid <= seg.hi)], [seg.lo, seg.hi])
            ph.append(x)
            ff.append(np.asarray(seg.f(x), dtype=float))
            lb.append(np.full(x.shape, bi))
            lp.append(np.array([_closest_piece(br, v, si) for v in x], dtype=int))
    P, F, LB, LP = map(np.concatenate, (ph, ff, lb, lp))
    o = np.lexsort((F, P))
    P, F, LB, LP = P[o], F[o], LB[o], LP[o]
    first = np.ones(P.shape, bool)
    first[1:] = P[1:] != P[:-1]  # one point per phi: the lowest f
    P, F, LB, LP = P[first], F[first], LB[first], LP[first]
    hull = _hull_lower(P, F)
    groups, slopes = [], {}
    for pos, k in enumerate(hull):
        lab = (int(LB[k]), int(LP[k]))
        if lab[1] < 0:
            warns.append(f"hull vertex at phi={P[k]:.6g} on concave-only segment of {names[lab[0]]!r} ignored")
            continue
        if groups and groups[-1] == lab:
            continue
        if groups:
            prev = hull[pos - 1]
            slopes[(groups[-1], lab)] = (F[k] - F[prev]) / (P[k] - P[prev])
        groups.append(lab)

    bridged = {}

    def bridge(gl, gr):
        if (gl, gr) not in bridged:
            bridged[(gl, gr)] = common_tangent(branches[gl[0]], branches[gr[0]], pieces=(gl[1], gr[1]),
                                               ordered=True, mu_guess=slopes.get((gl, gr)), n_scan=n_scan,
                                               atol=atol)
        return bridged[(gl, gr)]

    def lowest_dip(t):
        """(dip, branch, phi) of the branch lying furthest below t's line."""
        f0 = t.omega + t.mu * t.phi_a
        best = (math.inf, None, None)
        for bi, br in enumerate(branches):
            d, x = _below_line(br, t.mu, t.phi_a, f0)
            if d < best[0]:
                best = (d, bi, x)
        return best

    def dedup(gs):
        return [g for k, g in enumerate(gs) if k == 0 or g != gs[k - 1]]

    limit = max_iter if max_iter is not None else 4 * len(groups) + 4 * len(branches) + 10
    reported = set()
    for _ in range(limit):
        groups = dedup(groups)
        bridges = [bridge(groups[k], groups[k + 1]) for k in range(len(groups) - 1)]
        crossed = next((k for k in range(1, len(groups) - 1)
                        if bridges[k - 1].ok and bridges[k].ok and bridges[k - 1].phi_b > bridges[k].phi_a), None)
        if crossed is not None:
            del groups[crossed]
            continue
        inserted = False
        for k, t in enumerate(bridges):
            if not t.ok:
                continue
            d, bi, x = lowest_dip(t)
            if d >= -atol:
                continue
            seg = int(np.argmin([0 if s.lo <= x <= s.hi else 1 for s in branches[bi].segments]))
            lab = (bi, _closest_piece(branches[bi], x, seg))
            if lab[1] < 0 or lab in (groups[k], groups[k + 1]):
                if (k, lab) not in reported:
                    reported.add((k, lab))
                    warns.append(f"{names[bi]!r} dips {d:.3g} below the {t.a}-{t.b} tangent at phi={x:.6g} but "
                                 "cannot be inserted (same piece as a neighbour or concave)")
                continue
            groups.insert(k + 1, lab)
            inserted = True
            break
        if not inserted:
            break
    else:
        warns.append("envelope refinement did not converge; result may be inconsistent")
    groups = dedup(groups)
    bridges = [bridge(groups[k], groups[k + 1]) for k in range(len(groups) - 1)]
    bridges = [replace(t, stable=True) if t.ok else t for t in bridges]
    windows = []
    for k, (bi, pi) in enumerate(groups):
        p = branches[bi].pieces[pi]
        lo = bridges[k - 1].phi_b if k > 0 and bridges[k - 1].ok else p.lo
        hi = bridges[k].phi_a if k < len(bridges) and bridges[k].ok else p.hi
        windows.append(Window(names[bi], bi, pi, float(lo), float(hi)))
    for t in bridges:
        if not t.ok:
            warns.append(f"hull bridge {t.a}-{t.b} has no tangent ({t.status}: {t.message}); "
                         "the envelope there is domain-limited")
        elif t.across_gap:
            warns.append(f"bridge {t.a}-{t.b} spans a NaN gap of {t.a!r} (phi {t.phi_a:.6g} to {t.phi_b:.6g})")
    keys = []
    for t in bridges:
        if t.ok:
            keys += [(t.a, t.piece_a, t.b, t.piece_b, t.mu), (t.b, t.piece_b, t.a, t.piece_a, t.mu)]

    def on_envelope(c):
        return any((c.a, c.piece_a, c.b, c.piece_b) == s[:4] and abs(c.mu - s[4]) <= 1e-9 * (1.0 + abs(c.mu))
                   for s in keys)

    meta = []
    for i in range(len(branches)):
        for j in range(i, len(branches)):
            if i == j and len(branches[i].pieces) < 2:
                continue
            res = common_tangent(branches[i], branches[j], n_scan=n_scan, atol=atol)
            for c in (res.candidates or ((res,) if res.ok else ())):
                if on_envelope(c):
                    continue
                d, bi, _ = lowest_dip(c)
                if d < -atol:
                    meta.append(replace(c, stable=False, below=names[bi], candidates=()))
                else:
                    meta.append(replace(c, stable=False, candidates=()))
                    warns.append(f"tangent {c.a}-{c.b} at mu={c.mu:.6g} supports every branch but is not on the "
                                 "sampled hull; refine phi_grid")
    meta.sort(key=lambda c: c.mu)
    hv = np.array(hull, dtype=int)
    return Envelope(temp, names, tuple(windows), tuple(bridges), tuple(meta), tuple(warns), P[hv], F[hv])


# ---------------------------------------------------------------------------- Maxwell lines
@dataclass(frozen=True)
class TripleBracket:
    """Triple point L-X-R between consecutive temperatures: X's window exists at
    `temp_open` (width `width_open`) and is gone at `temp_closed`, where
    `width_closed` is the signed width from the metastable L-X and X-R
    tangents (negative) when both exist, and `temp_estimate` the linear zero."""
    phases: tuple
    temp_lo: float
    temp_hi: float
    temp_open: float
    temp_closed: float
    width_open: float
    width_closed: float | None = None
    temp_estimate: float | None = None


# column order of the rows / metastable tables and their CSV
_ROW_KEYS = ("temp", "a", "b", "phi_a", "phi_b", "mu", "omega", "q0star_a", "q0star_b", "status", "same_branch",
             "across_gap", "stable", "below", "residual")


def _row(temp, t: TangentResult) -> dict:
    return {k: (temp if k == "temp" else (t.residual if t.ok else math.nan) if k == "residual" else getattr(t, k))
            for k in _ROW_KEYS}


@dataclass(frozen=True)
class MaxwellLines:
    """Coexistence over temperatures: `rows` (stable tangents, by temperature
    then density), `metastable` (same columns), `windows` (temp, name,
    phi_lo, phi_hi), `triple_points` (TripleBracket list), `temps`."""
    rows: tuple
    metastable: tuple
    windows: tuple
    triple_points: tuple
    temps: tuple

    def columns(self, which: str = "rows") -> dict:
        """Column-oriented copy of `rows`, `metastable` or `windows`."""
        data = getattr(self, which)
        if data:
            keys = list(data[0].keys())
        else:
            keys = ["temp", "name", "phi_lo", "phi_hi"] if which == "windows" else list(_ROW_KEYS)
        return {k: [r[k] for r in data] for k in keys}

    def write_csv(self, dest, which: str = "rows") -> None:
        """`rows`, `metastable` or `windows` as CSV: a header, floats as repr,
        None as an empty field; `dest` is a path (ASCII, LF) or an open text file."""
        keys = list(self.columns(which))

        def cell(v):
            return "" if v is None else repr(v) if isinstance(v, float) else v

        def emit(fh):
            w = csv.writer(fh, lineterminator="\n")
            w.writerow(keys)
            for r in getattr(self, which):
                w.writerow([cell(r[k]) for k in keys])

        if hasattr(dest, "write"):
            emit(dest)
        else:
            with open(dest, "w", newline="", encoding="ascii") as fh:
                emit(fh)


def _window_width(env: Envelope, L: str, X: str, R: str, mu_ref):
    """Signed width of X's window between L and R from env's metastable L-X and
    X-R tangents (X-side densities), or None."""
    def x_side(c, left, right):
        if c.a == left and c.b == right and c.phi_a < c.phi_b:
            return c.phi_b if right == X else c.phi_a
        if c.a == right and c.b == left and c.phi_b < c.phi_a:
            return c.phi_a if right == X else c.phi_b
        return None

    lx, xr = [], []
    for c in env.metastable:
        if not c.ok:
            continue
        v = x_side(c, L, X)
        if v is not None:
            lx.append((c.mu, v))
        v = x_side(c, X, R)
        if v is not None:
            xr.append((c.mu, v))
    if not lx or not xr:
        return None
    if mu_ref is None:
        return xr[0][1] - lx[0][1]
    near = min(xr, key=lambda t: abs(t[0] - mu_ref)), min(lx, key=lambda t: abs(t[0] - mu_ref))
    return near[0][1] - near[1][1]


def maxwell_lines(per_temperature_results: Mapping | Iterable[Envelope]) -> MaxwellLines:
    """Envelopes over temperatures ({temp: Envelope}, or envelopes built with
    `temp`) as tables, with triple points bracketed between consecutive
    temperatures in both directions."""
    if isinstance(per_temperature_results, Mapping):
        items = [(float(t), e) for t, e in per_temperature_results.items()]
    else:
        items = []
        for e in per_temperature_results:
            if e.temp is None:
                raise ValueError("maxwell_lines: envelope without temp; pass a mapping")
            items.append((float(e.temp), e))
    items.sort(key=lambda te: te[0])
    rows, meta, wins = [], [], []
    for t, e in items:
        rows += [_row(t, c) for c in e.coexistence]
        meta += [_row(t, c) for c in e.metastable]
        wins += [{"temp": t, "name": w.name, "phi_lo": w.phi_lo, "phi_hi": w.phi_hi} for w in e.windows]
    triples = []
    for (t1, e1), (t2, e2) in zip(items[:-1], items[1:]):
        for (ta, ea), (tb, eb) in (((t1, e1), (t2, e2)), ((t2, e2), (t1, e1))):
            sa, sb = [w.name for w in ea.windows], [w.name for w in eb.windows]
            for k in range(1, len(sa) - 1):
                L, X, R = sa[k - 1], sa[k], sa[k + 1]
                adjacent = [m for m in range(len(sb) - 1) if sb[m] == L and sb[m + 1] == R]
                still = any(sb[m - 1:m + 2] == [L, X, R] for m in range(1, len(sb) - 1))
                if not adjacent or still:
                    continue
                w_open = ea.windows[k].width
                c = eb.coexistence[adjacent[0]]
                w_closed = _window_width(eb, L, X, R, c.mu if c.ok else None)
                est = None
                if w_closed is not None and w_closed < 0.0 < w_open:
                    est = ta + (tb - ta) * w_open / (w_open - w_closed)
                triples.append(TripleBracket((L, X, R), min(ta, tb), max(ta, tb), ta, tb, float(w_open), w_closed,
                                             est))
    return MaxwellLines(tuple(rows), tuple(meta), tuple(wins), tuple(triples), tuple(t for t, _ in items))
