"""Coexistence of free-energy branches of one conserved density.

The drop-in counterpart of the reference's phase_diagram/common_tangent.py: the
same public names (Segment, Piece, Branch, TangentResult, Window, Envelope,
TripleBracket, MaxwellLines, analytic_branch, tabulated_branch,
common_tangent, lower_envelope, maxwell_lines), result fields, status names
and CSV columns, so tables written by either read the same.  The
construction is this module's own.

A branch f(phi) is the free energy per site of one phase as a function of the
mean density: analytic (callables, e.g. a uniform state) or tabulated (the
relaxed fmin(phi) of a q0* scan, a not-a-knot cubic spline per NaN-free run).
It is cut into convex pieces (f'' > 0, so mu = f' is increasing and has an
inverse) and concave intervals.  Two phases coexist where one line touches
both: equal mu and equal grand potential omega = f - mu phi.

Tangents.  For convex pieces P (of branch a) and Q (of b) with overlapping mu
ranges, let x run over the densities of P whose slope Q also takes, y(x) the
density on Q with mu_b(y) = mu_a(x), and

    h(x) = f_a(x) - f_b(y(x)) - mu_a(x) (x - y(x)),

the height of P's tangent line at x above Q's parallel tangent line.  A root
of h is a common tangent; h' = -f_a''(x) (x - y), so h is monotone while the
two contact points keep their order.  h is sampled at `n_scan` densities,
every sign change is refined by Brent's method, and y(x) is found by a
safeguarded Newton iteration on mu_b (f'' where the branch has it).

Envelope.  The lower convex hull of the convex pieces of all branches is
built exactly by gift wrapping, from the leftmost density: on the current
piece, the hull leaves at the tangent of smallest slope towards any piece to
the right (tangent slopes grow along a convex piece, so this is the first
contact of a line rotated about the curve); with no such tangent it runs to
the piece's end and jumps by the chord of smallest slope.  Every other
pairwise tangent is metastable and names the branch that dips deepest below
its line.  No sampling grid decides the hull; `phi_grid` (or `n` points) only
tabulates the envelope in `Envelope.hull_phi/hull_f`.

Statuses of a pair without a tangent (the reference's contract):
``no_overlap`` (mu ranges disjoint, concave parts included), ``no_root``
(convex pieces overlap in mu but never touch one line; also a self pair of a
branch with one convex piece) and ``non_convex`` (no convex piece, or
overlap through concave parts only).

phi: mean density; mu = df/dphi; omega = f - mu phi; `temp`: the model's
temperature-like parameter (epsilon for the PFC models).
"""
from __future__ import annotations

import csv
import math
from collections.abc import Callable, Iterable, Mapping, Sequence
from dataclasses import asdict, dataclass, field, replace

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.optimize import brentq

__all__ = ["STATUSES", "Segment", "Piece", "Branch", "TangentResult", "Window", "Envelope", "TripleBracket",
           "MaxwellLines", "analytic_branch", "tabulated_branch", "common_tangent", "lower_envelope",
           "maxwell_lines"]

STATUSES = ("ok", "no_overlap", "no_root", "non_convex")

_EPS = np.finfo(float).eps


def _scalar(fn, x):
    """fn at one density, as a float (callables are vectorised)."""
    return float(np.asarray(fn(np.array([float(x)])), dtype=float).reshape(-1)[0])


# ---------------------------------------------------------------------------
# branch data
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Segment:
    """A density interval [lo, hi] with one representation: vectorised `f`,
    `mu` (= f'), optional `d2f` and `q0star`; `nodes` are a table's densities."""
    lo: float
    hi: float
    f: Callable
    mu: Callable
    d2f: Callable | None = None
    q0star: Callable | None = None
    nodes: np.ndarray | None = field(default=None, repr=False, compare=False)


@dataclass(frozen=True)
class Piece:
    """A convex part of segment `segment`: mu rises from mu_lo to mu_hi on [lo, hi]."""
    segment: int
    lo: float
    hi: float
    mu_lo: float
    mu_hi: float


@dataclass(frozen=True)
class Branch:
    """f(phi) of one phase: `segments` (one per NaN-free run of a table),
    their convex `pieces` and concave intervals `nonconvex` in density order,
    `gaps` between segments and `dropped` table densities (runs too short)."""
    name: str
    kind: str
    segments: tuple
    pieces: tuple
    nonconvex: tuple
    gaps: tuple = ()
    dropped: tuple = ()
    _dense: tuple = field(default=(), repr=False, compare=False)  # (phi, f, mu) samples

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
        """(min, max) of mu over the sampled segments, concave parts included."""
        mu = self._dense[2]
        return float(mu.min()), float(mu.max())

    def _apply(self, which, phi):
        x = np.asarray(phi, dtype=float)
        flat = np.atleast_1d(x).astype(float).ravel()
        out = np.full(flat.shape, np.nan)
        for seg in self.segments:
            fn = getattr(seg, which)
            inside = (flat >= seg.lo) & (flat <= seg.hi)
            if fn is not None and inside.any():
                out[inside] = np.asarray(fn(flat[inside]), dtype=float)
        return float(out[0]) if x.ndim == 0 else out.reshape(x.shape)

    def f(self, phi):
        """f(phi), NaN outside the segments."""
        return self._apply("f", phi)

    def mu(self, phi):
        """mu(phi) = f'(phi), NaN outside the segments."""
        return self._apply("mu", phi)

    def q0star(self, phi):
        """q0*(phi) where the branch carries it, NaN elsewhere."""
        return self._apply("q0star", phi)

    def piece_at(self, phi: float):
        """Index of the convex piece holding `phi`, or None."""
        hits = [k for k, p in enumerate(self.pieces) if p.lo <= phi <= p.hi]
        return hits[0] if hits else None


def _dense_samples(segments, n):
    """(phi, f, mu) on `n` points per segment plus its table nodes."""
    xs = []
    for seg in segments:
        x = np.linspace(seg.lo, seg.hi, n)
        xs.append(x if seg.nodes is None else np.union1d(x, seg.nodes))
    phi = np.concatenate(xs)
    f = np.concatenate([np.asarray(s.f(x), dtype=float) for s, x in zip(segments, xs)])
    mu = np.concatenate([np.asarray(s.mu(x), dtype=float) for s, x in zip(segments, xs)])
    return phi, f, mu


def _split(segments, cuts, convex_at):
    """Cut each segment at its `cuts`; the intervals where `convex_at(seg, u,
    v)` holds become pieces (touching ones merged), the rest concave."""
    pieces, concave = [], []
    for k, seg in enumerate(segments):
        edges = np.unique(np.concatenate([[seg.lo, seg.hi], np.clip(np.asarray(cuts[k], float), seg.lo, seg.hi)]))
        span = max(seg.hi - seg.lo, 1e-300)
        for u, v in zip(edges[:-1].tolist(), edges[1:].tolist()):
            if v - u <= 1e-12 * span:
                continue
            if convex_at(seg, u, v):
                if pieces and pieces[-1].segment == k and pieces[-1].hi == u:
                    pieces[-1] = replace(pieces[-1], hi=v, mu_hi=_scalar(seg.mu, v))
                else:
                    pieces.append(Piece(k, u, v, _scalar(seg.mu, u), _scalar(seg.mu, v)))
            elif concave and concave[-1][1] == u:
                concave[-1] = (concave[-1][0], v)
            else:
                concave.append((u, v))
    return tuple(pieces), tuple(concave)


def _sign_change_roots(g, x, vals, xtol):
    """Roots of g at the sign changes of `vals` = g(x) on the sorted grid x."""
    out = []
    for i in np.flatnonzero(np.sign(vals[:-1]) * np.sign(vals[1:]) < 0):
        out.append(brentq(g, x[i], x[i + 1], xtol=xtol, rtol=4 * _EPS))
    out.extend(x[np.flatnonzero(vals == 0.0)].tolist())
    return sorted(out)


def analytic_branch(name, f, mu, lo, hi, *, d2f=None, q0star=None, n_sample=2049):
    """Branch of vectorised callables on [lo, hi] (`mu` = df/dphi).  The
    concave intervals are where `d2f` < 0, or, without it, where the sampled
    mu falls; each boundary is refined to a root of f'' (of a centred
    difference of mu without `d2f`)."""
    lo, hi = float(lo), float(hi)
    if not hi > lo:
        raise ValueError(f"analytic_branch {name!r}: need lo < hi, got {lo}, {hi}")
    seg = Segment(lo, hi, f, mu, d2f, q0star)
    x = np.linspace(lo, hi, n_sample)
    if d2f is not None:
        g = lambda p: _scalar(d2f, p)  # noqa: E731
        cuts = _sign_change_roots(g, x, np.asarray(d2f(x), dtype=float), 1e-15)
        convex_at = lambda s, u, v: _scalar(s.d2f, 0.5 * (u + v)) > 0.0  # noqa: E731
    else:
        m = np.asarray(mu(x), dtype=float)
        step = 1e-6 * max(1.0, abs(lo), abs(hi))

        def slope(p):  # centred difference of mu, one-sided at the ends
            a, b = max(lo, p - step), min(hi, p + step)
            return (_scalar(mu, b) - _scalar(mu, a)) / (b - a)

        # a turning point of mu sits between samples where its increments
        # change sign; refine it as the root of mu's difference quotient
        inc = np.sign(np.diff(m))
        cuts = []
        for i in np.flatnonzero(inc[:-1] * inc[1:] < 0):
            a, b = x[i], x[i + 2]
            sa, sb = slope(a), slope(b)
            cuts.append(brentq(slope, a, b, xtol=1e-14, rtol=4 * _EPS) if sa * sb < 0 else x[i + 1])
        convex_at = lambda s, u, v: _scalar(s.mu, v) > _scalar(s.mu, u)  # noqa: E731
    pieces, concave = _split((seg,), [cuts], convex_at)
    return Branch(name, "analytic", (seg,), pieces, concave, _dense=_dense_samples((seg,), n_sample))


def _spline_inflections(spl):
    """Zeros of f'' of a cubic spline: f'' is linear between knots, so each
    sign change of its knot values has one zero, found by interpolation."""
    x = spl.x
    d2 = spl(x, 2)
    out = [float(x[i]) for i in np.flatnonzero(d2 == 0.0)]
    for i in np.flatnonzero(d2[:-1] * d2[1:] < 0):
        t = d2[i] / (d2[i] - d2[i + 1])
        out.append(float(x[i] + t * (x[i + 1] - x[i])))
    return sorted(out)


def tabulated_branch(name, phi, f, q0star=None, *, min_points=4, bc_type="not-a-knot", n_sample=1025):
    """Branch of a table f(phi) with NaN where a relaxation failed: each run
    of >= `min_points` consecutive finite entries (phi, f and q0star if
    given) becomes a segment with cubic splines of f (mu and f'' its
    derivatives) and q0*; shorter runs are listed in `dropped`, the spaces
    between segments in `gaps`."""
    phi = np.asarray(phi, dtype=float).ravel()
    f = np.asarray(f, dtype=float).ravel()
    q = None if q0star is None else np.asarray(q0star, dtype=float).ravel()
    if f.shape != phi.shape or (q is not None and q.shape != phi.shape):
        raise ValueError(f"tabulated_branch {name!r}: phi, f, q0star must have equal length")
    idx = np.argsort(phi, kind="stable")
    phi, f = phi[idx], f[idx]
    q = None if q is None else q[idx]
    fin = np.isfinite(phi)
    if np.any(np.diff(phi[fin]) == 0):
        raise ValueError(f"tabulated_branch {name!r}: repeated phi values")
    ok = fin & np.isfinite(f) & (np.isfinite(q) if q is not None else True)
    # runs of consecutive usable entries: boundaries where `ok` switches
    flips = np.flatnonzero(np.diff(np.concatenate([[0], ok.astype(int), [0]])))
    segments, cuts, dropped = [], [], []
    for s, e in zip(flips[0::2], flips[1::2]):
        if e - s < min_points:
            dropped += phi[s:e].tolist()
            continue
        x = phi[s:e]
        spl = CubicSpline(x, f[s:e], bc_type=bc_type)
        qs = None if q is None else CubicSpline(x, q[s:e], bc_type=bc_type)
        segments.append(Segment(float(x[0]), float(x[-1]), spl, spl.derivative(1), spl.derivative(2), qs, x))
        cuts.append(_spline_inflections(spl))
    if not segments:
        raise ValueError(f"tabulated_branch {name!r}: no run of >= {min_points} finite points")
    pieces, concave = _split(segments, cuts, lambda s, u, v: _scalar(s.d2f, 0.5 * (u + v)) > 0.0)
    gaps = tuple((a.hi, b.lo) for a, b in zip(segments[:-1], segments[1:]))
    return Branch(name, "tabulated", tuple(segments), pieces, concave, gaps, tuple(dropped),
                  _dense=_dense_samples(segments, n_sample))


# ---------------------------------------------------------------------------
# the pairwise tangent
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class TangentResult:
    """Common tangent of branches `a` and `b` (names), or the reason for none.

    phi_a, phi_b: contact densities on a and b (argument order; either may be
    the lower).  residual_mu_a/b = mu(phi) - mu, residual_omega = omega_a -
    omega_b.  piece_a/b index Branch.pieces.  `supporting`: neither branch of
    the pair lies below the line by more than atol.  `stable`/`below` are set
    by lower_envelope.  `candidates`: every tangent of the search."""
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
        del d["candidates"]
        return d


def _contact(br: Branch, k: int, m: float) -> float:
    """Density on convex piece k of `br` where mu = m (clamped to the piece):
    Newton steps with the branch's f'', kept inside a shrinking bracket and
    bisected when a step leaves it; Brent's method where f'' is not given."""
    p = br.pieces[k]
    if m <= p.mu_lo:
        return p.lo
    if m >= p.mu_hi:
        return p.hi
    seg = br.segments[p.segment]
    if seg.d2f is None:
        return brentq(lambda v: _scalar(seg.mu, v) - m, p.lo, p.hi, xtol=1e-15, rtol=4 * _EPS)
    lo, hi = p.lo, p.hi
    x = lo + (hi - lo) * (m - p.mu_lo) / (p.mu_hi - p.mu_lo)  # linear start
    for _ in range(100):
        gx = _scalar(seg.mu, x) - m
        if gx == 0.0:
            return x
        if gx < 0:
            lo = x
        else:
            hi = x
        d = _scalar(seg.d2f, x)
        nx = x - gx / d if d > 0 else 0.5 * (lo + hi)
        if not lo < nx < hi:
            nx = 0.5 * (lo + hi)
        if abs(nx - x) <= 2 * _EPS * max(1.0, abs(x)) or hi - lo <= 2 * _EPS * max(1.0, abs(x)):
            return nx
        x = nx
    return x


def _line_gap(br: Branch, m: float, x0: float, f0: float):
    """Lowest value of f - (f0 + m (phi - x0)) over the branch's dense
    samples, and where it is."""
    phi, f, _ = br._dense
    d = f - (f0 + m * (phi - x0))
    i = int(np.argmin(d))
    return float(d[i]), float(phi[i])


def _touching(a: Branch, i: int, b: Branch, j: int, n_scan: int):
    """Densities x on piece i of `a` where one line touches piece j of `b`."""
    pa, pb = a.pieces[i], b.pieces[j]
    m_lo, m_hi = max(pa.mu_lo, pb.mu_lo), min(pa.mu_hi, pb.mu_hi)
    if not m_lo < m_hi:
        return None
    sa, sb = a.segments[pa.segment], b.segments[pb.segment]
    x_lo, x_hi = _contact(a, i, m_lo), _contact(a, i, m_hi)

    def h(x):
        m = _scalar(sa.mu, x)
        y = _contact(b, j, m)
        return _scalar(sa.f, x) - _scalar(sb.f, y) - m * (x - y)

    if not x_lo < x_hi:
        return []
    xs = np.linspace(x_lo, x_hi, max(int(n_scan), 2))
    vals = np.array([h(x) for x in xs])
    roots = _sign_change_roots(h, xs, vals, 1e-15)
    uniq = []
    for r in roots:
        if not uniq or r - uniq[-1] > 1e-12 * (1.0 + abs(r)):
            uniq.append(r)
    return uniq


def _tangent_at(a, i, b, j, x, atol):
    """The TangentResult of the contact at density x of piece i of `a`."""
    same = a is b
    sa, sb = a.segments[a.pieces[i].segment], b.segments[b.pieces[j].segment]
    m = _scalar(sa.mu, x)
    y = _contact(b, j, m)
    fa, fb = _scalar(sa.f, x), _scalar(sb.f, y)
    om_a, om_b = fa - m * x, fb - m * y
    low = min(_line_gap(a, m, x, fa)[0], _line_gap(b, m, x, fa)[0])
    return TangentResult(a.name, b.name, "ok", float(x), float(y), float(m), float(om_a),
                         None if sa.q0star is None else _scalar(sa.q0star, x),
                         None if sb.q0star is None else _scalar(sb.q0star, y),
                         _scalar(sa.mu, x) - m, _scalar(sb.mu, y) - m, float(om_a - om_b), i, j, same,
                         bool(same and a.pieces[i].segment != b.pieces[j].segment), bool(low >= -atol))


def common_tangent(a: Branch, b: Branch, *, pieces=None, ordered=False, mu_guess=None, n_scan=33, atol=1e-10):
    """Common tangent(s) of branches a and b over every pair of convex pieces
    (distinct pieces when a is b: a self-tangent across a concave interval).
    `pieces` = (i, j) restricts the search to one pair; `ordered` keeps
    tangents with phi_a < phi_b; the primary result is a supporting tangent
    before a non-supporting one, then the one nearest `mu_guess` (else the
    lowest mu), and `.candidates` holds all.  A pair without a tangent gives
    a non-ok status (see the module docstring), never an exception."""
    same = a is b
    empty = [br.name for br in (a, b) if not br.pieces]
    if empty:
        return TangentResult(a.name, b.name, "non_convex", same_branch=same,
                             message=f"branch {empty[0]!r} has no convex piece")
    if pieces is not None:
        pairs = [tuple(pieces)]
    else:
        pairs = [(i, j) for i in range(len(a.pieces)) for j in range(len(b.pieces)) if not same or j > i]
    found, overlapped = [], False
    for i, j in pairs:
        xs = _touching(a, i, b, j, n_scan)
        if xs is None:
            continue
        overlapped = True
        found += [_tangent_at(a, i, b, j, x, atol) for x in xs]
    if ordered:
        found = [t for t in found if t.phi_a < t.phi_b]
    if found:
        rank = (lambda t: (not t.supporting, abs(t.mu - mu_guess))) if mu_guess is not None else \
            (lambda t: (not t.supporting, t.mu))
        found.sort(key=rank)
        best = replace(found[0], candidates=tuple(found) if pieces is None else ())
        if len(found) > 1:
            best = replace(best, message=f"{len(found)} tangents; primary: supporting first, then "
                                         f"{'nearest mu_guess' if mu_guess is not None else 'lowest mu'}")
        return best
    if same and len(a.pieces) < 2:
        return TangentResult(a.name, b.name, "no_root", same_branch=True,
                             message="one convex piece: no tangent of the branch with itself")
    (a0, a1), (b0, b1) = a.mu_range, b.mu_range
    if max(a0, b0) >= min(a1, b1):
        return TangentResult(a.name, b.name, "no_overlap", same_branch=same,
                             message=f"mu ranges [{a0:.6g}, {a1:.6g}] and [{b0:.6g}, {b1:.6g}] do not overlap")
    if overlapped:
        return TangentResult(a.name, b.name, "no_root", same_branch=same,
                             message="convex pieces share slopes but no line touches both"
                                     + (" with phi_a < phi_b" if ordered else ""))
    return TangentResult(a.name, b.name, "non_convex", same_branch=same,
                         message="the mu ranges meet only on concave parts")


# ---------------------------------------------------------------------------
# the lower envelope
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Window:
    """Densities [phi_lo, phi_hi] where phase `name` (branch index `branch`,
    convex piece `piece`) alone is on the envelope."""
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
    windows[k+1] in density order; `metastable`: the other tangents (stable
    False, `below` the branch under the line); hull_phi/hull_f: the envelope
    tabulated for plots."""
    temp: float | None
    names: tuple
    windows: tuple
    coexistence: tuple
    metastable: tuple
    warnings: tuple
    hull_phi: np.ndarray = field(repr=False, compare=False, default=None)
    hull_f: np.ndarray = field(repr=False, compare=False, default=None)


def _deepest(branches, res):
    """(depth, branch index, phi) of the branch lying lowest under the line of `res`."""
    f0 = res.omega + res.mu * res.phi_a
    worst = (math.inf, None, None)
    for bi, br in enumerate(branches):
        d, x = _line_gap(br, res.mu, res.phi_a, f0)
        if d < worst[0]:
            worst = (d, bi, x)
    return worst


def _chord(branches, labels, x0, f0, exclude):
    """The chord of smallest slope from (x0, f0) to a convex piece lying
    right of x0: (label, contact density)."""
    best = (math.inf, None, None)
    for lab in labels:
        if lab == exclude:
            continue
        br = branches[lab[0]]
        p = br.pieces[lab[1]]
        if p.hi <= x0:
            continue
        seg = br.segments[p.segment]
        x = np.linspace(max(p.lo, x0), p.hi, 257)[1:] if p.lo <= x0 else np.linspace(p.lo, p.hi, 257)
        s = (np.asarray(seg.f(x), dtype=float) - f0) / (x - x0)
        k = int(np.argmin(s))
        if s[k] < best[0]:
            best = (float(s[k]), lab, float(x[k]))
    return best[1], best[2]


def lower_envelope(branches: Sequence, phi_grid=None, *, temp=None, n=4001, n_scan=33, atol=1e-10, max_iter=None):
    """Stable coexistence and single-phase windows of all `branches` at one
    temperature (gift wrapping over their convex pieces, see the module
    docstring); every other pairwise tangent is returned as metastable.
    `phi_grid` (default `n` points over all domains) tabulates the envelope.
    `max_iter` bounds the hull's steps (default: enough for every piece)."""
    branches = list(branches)
    if not branches:
        raise ValueError("lower_envelope: no branches")
    names = tuple(b.name for b in branches)
    if len(set(names)) != len(names):
        raise ValueError(f"lower_envelope: branch names must be unique, got {names}")
    warns = []
    labels = [(bi, pi) for bi, br in enumerate(branches) for pi in range(len(br.pieces))]
    if not labels:
        raise ValueError("lower_envelope: no branch has a convex piece")
    memo = {}

    def tangents(la, lb):
        key = (la, lb)
        if key not in memo:
            a, b = branches[la[0]], branches[lb[0]]
            xs = _touching(a, la[1], b, lb[1], n_scan) or []
            memo[key] = [_tangent_at(a, la[1], b, lb[1], x, atol) for x in xs]
        return memo[key]

    def start_of(lab):
        p = branches[lab[0]].pieces[lab[1]]
        return p.lo, _scalar(branches[lab[0]].segments[p.segment].f, p.lo)

    # leftmost contact: the lowest f at the smallest density
    cur = min(labels, key=lambda lab: start_of(lab))
    x_in = start_of(cur)[0]
    seq, bridges, entries = [cur], [], [x_in]
    steps = max_iter if max_iter is not None else 4 * len(labels) + 8
    for _ in range(steps):
        best, nxt = None, None
        for other in labels:
            if other == cur:
                continue
            for t in tangents(cur, other):
                if t.phi_a >= x_in - 1e-12 * (1 + abs(x_in)) and t.phi_b > t.phi_a and (best is None or t.mu < best.mu):
                    best, nxt = t, other
        if best is not None:
            bridges.append(replace(best, stable=True))
            cur, x_in = nxt, best.phi_b
        else:
            p = branches[cur[0]].pieces[cur[1]]
            x0, f0 = p.hi, _scalar(branches[cur[0]].segments[p.segment].f, p.hi)
            nxt, xn = _chord(branches, labels, x0, f0, cur)
            if nxt is None:
                break
            a, b = branches[cur[0]], branches[nxt[0]]
            res = common_tangent(a, b, pieces=(cur[1], nxt[1]), ordered=True, n_scan=n_scan, atol=atol)
            bridges.append(res)
            warns.append(f"hull bridge {res.a}-{res.b} has no tangent ({res.status}: {res.message}); "
                         "the envelope there is domain-limited")
            cur, x_in = nxt, xn
        seq.append(cur)
        entries.append(x_in)
    else:
        warns.append("envelope: the hull did not reach the right end within max_iter steps")
    windows = []
    for k, lab in enumerate(seq):
        p = branches[lab[0]].pieces[lab[1]]
        lo = bridges[k - 1].phi_b if k > 0 and bridges[k - 1].ok else (entries[k] if k > 0 else p.lo)
        hi = bridges[k].phi_a if k < len(bridges) and bridges[k].ok else p.hi
        windows.append(Window(names[lab[0]], lab[0], lab[1], float(lo), float(hi)))
    for br in bridges:
        if br.ok:
            depth, bi, x = _deepest(branches, br)
            if depth < -atol:
                warns.append(f"{names[bi]!r} dips {depth:.3g} below the {br.a}-{br.b} tangent at phi={x:.6g} "
                             "(on a concave part or a piece without a tangent)")
            if br.across_gap:
                warns.append(f"bridge {br.a}-{br.b} spans a NaN gap of {br.a!r} (phi {br.phi_a:.6g} to {br.phi_b:.6g})")
    on_hull = {(b.a, b.piece_a, b.b, b.piece_b, round(b.mu, 9)) for b in bridges if b.ok}
    meta = []
    for ia, la in enumerate(labels):
        for lb in labels[ia + 1:]:
            for t in tangents(la, lb):
                key = (t.a, t.piece_a, t.b, t.piece_b, round(t.mu, 9))
                rev = (t.b, t.piece_b, t.a, t.piece_a, round(t.mu, 9))
                if key in on_hull or rev in on_hull:
                    continue
                depth, bi, _ = _deepest(branches, t)
                if depth < -atol:
                    meta.append(replace(t, stable=False, below=names[bi]))
                else:
                    meta.append(replace(t, stable=False))
                    warns.append(f"tangent {t.a}-{t.b} at mu={t.mu:.6g} supports every branch but is not on the hull")
    meta.sort(key=lambda t: t.mu)
    grid = (np.linspace(min(b.lo for b in branches), max(b.hi for b in branches), n) if phi_grid is None
            else np.asarray(phi_grid, dtype=float))
    hull_f = np.full(grid.shape, np.nan)
    for k, w in enumerate(windows):
        inside = (grid >= w.phi_lo) & (grid <= w.phi_hi)
        hull_f[inside] = branches[w.branch].f(grid[inside])
        if k < len(bridges) and bridges[k].ok:
            t = bridges[k]
            gap = (grid > t.phi_a) & (grid < t.phi_b)
            hull_f[gap] = t.omega + t.mu * grid[gap]
    return Envelope(temp, names, tuple(windows), tuple(bridges), tuple(meta), tuple(warns), grid, hull_f)


# ---------------------------------------------------------------------------
# Maxwell lines over temperatures
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class TripleBracket:
    """Phases (L, X, R): X's window between L and R exists at `temp_open` and
    has closed (L and R coexist) at the neighbouring `temp_closed`.
    `width_closed` is X's signed width there from its metastable tangents
    with L and R (negative); with both widths, `temp_estimate` is the linear
    zero of the width."""
    phases: tuple
    temp_lo: float
    temp_hi: float
    temp_open: float
    temp_closed: float
    width_open: float
    width_closed: float | None = None
    temp_estimate: float | None = None


_ROW_KEYS = ("temp", "a", "b", "phi_a", "phi_b", "mu", "omega", "q0star_a", "q0star_b", "status", "same_branch",
             "across_gap", "stable", "below", "residual")
_WINDOW_KEYS = ("temp", "name", "phi_lo", "phi_hi")


def _as_row(temp, t: TangentResult) -> dict:
    row = {k: getattr(t, k) for k in _ROW_KEYS[1:-1]}
    return {"temp": temp, **row, "residual": t.residual if t.ok else math.nan}


@dataclass(frozen=True)
class MaxwellLines:
    """Coexistence over temperatures: `rows` (stable tangents, by
    temperature then density), `metastable`, `windows` (dicts) and
    `triple_points`."""
    rows: tuple
    metastable: tuple
    windows: tuple
    triple_points: tuple
    temps: tuple

    def columns(self, which: str = "rows") -> dict:
        """Column-wise copy of `rows`, `metastable` or `windows`."""
        data = getattr(self, which)
        keys = list(data[0]) if data else list(_WINDOW_KEYS if which == "windows" else _ROW_KEYS)
        return {k: [r[k] for r in data] for k in keys}

    def write_csv(self, dest, which: str = "rows") -> None:
        """`rows`, `metastable` or `windows` as CSV (floats as repr, None as
        ''); `dest` is a path (ASCII, LF) or an open text file."""
        keys = list(self.columns(which))

        def cell(v):
            return "" if v is None else (repr(v) if isinstance(v, float) else v)

        def write(fh):
            out = csv.writer(fh, lineterminator="\n")
            out.writerow(keys)
            out.writerows([cell(r[k]) for k in keys] for r in getattr(self, which))

        if hasattr(dest, "write"):
            write(dest)
        else:
            with open(dest, "w", newline="", encoding="ascii") as fh:
                write(fh)


def _window_edge(env: Envelope, left: str, right: str, side: str, mu_ref):
    """X's edge from env's metastable tangents of X with its neighbour:
    `side` "lo": the X density of an (L, X) tangent with L below X;
    "hi": the X density of an (X, R) tangent with X below R.  The tangent
    nearest `mu_ref` (else the first) when several exist."""
    got = []
    for t in env.metastable:
        if not t.ok:
            continue
        for (p, q, xp, xq) in ((t.a, t.b, t.phi_a, t.phi_b), (t.b, t.a, t.phi_b, t.phi_a)):
            if (p, q) == (left, right) and xp < xq:
                got.append((t.mu, xq if side == "lo" else xp))
    if not got:
        return None
    if mu_ref is None:
        return got[0][1]
    return min(got, key=lambda g: abs(g[0] - mu_ref))[1]


def maxwell_lines(per_temperature_results) -> MaxwellLines:
    """Collect lower_envelope results over temperatures ({temp: Envelope} or
    envelopes built with `temp`) and bracket triple points between
    consecutive temperatures."""
    if isinstance(per_temperature_results, Mapping):
        items = [(float(t), e) for t, e in per_temperature_results.items()]
    else:
        items = []
        for e in per_temperature_results:
            if e.temp is None:
                raise ValueError("maxwell_lines: envelope without temp; pass a mapping")
            items.append((float(e.temp), e))
    items.sort(key=lambda it: it[0])
    rows = tuple(_as_row(t, c) for t, e in items for c in e.coexistence)
    meta = tuple(_as_row(t, c) for t, e in items for c in e.metastable)
    wins = tuple({"temp": t, "name": w.name, "phi_lo": w.phi_lo, "phi_hi": w.phi_hi} for t, e in items
                 for w in e.windows)
    triples = []
    for (t1, e1), (t2, e2) in zip(items, items[1:]):
        for (t_open, eo), (t_closed, ec) in (((t1, e1), (t2, e2)), ((t2, e2), (t1, e1))):
            seq_o = [w.name for w in eo.windows]
            seq_c = [w.name for w in ec.windows]
            pairs_c = {(seq_c[m], seq_c[m + 1]): m for m in range(len(seq_c) - 1)}
            triples_c = {tuple(seq_c[m:m + 3]) for m in range(len(seq_c) - 2)}
            for k in range(1, len(seq_o) - 1):
                L, X, R = seq_o[k - 1:k + 2]
                if (L, R) not in pairs_c or (L, X, R) in triples_c:
                    continue
                bridge = ec.coexistence[pairs_c[(L, R)]]
                mu_ref = bridge.mu if bridge.ok else None
                lo = _window_edge(ec, L, X, "lo", mu_ref)
                hi = _window_edge(ec, X, R, "hi", mu_ref)
                w_open = float(eo.windows[k].width)
                w_closed = None if lo is None or hi is None else hi - lo
                est = None
                if w_closed is not None and w_closed < 0.0 < w_open:
                    est = t_open + (t_closed - t_open) * w_open / (w_open - w_closed)
                triples.append(TripleBracket((L, X, R), min(t_open, t_closed), max(t_open, t_closed), t_open,
                                             t_closed, w_open, w_closed, est))
    return MaxwellLines(rows, meta, wins, tuple(triples), tuple(t for t, _ in items))
