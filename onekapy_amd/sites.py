"""Synthetic benchmark sites (seeded), in the reference's input vocabulary.

Each site is what ``create_stochastic_capturezone`` takes (wells with
discharge distributions, aquifer distributions, observations) plus the run
sizes.  Observed heads come from a "true" regional potential
Phi = A dx^2 + B dy^2 + C dx dy + D dx + E dy + F (origin at the target) plus
the wells at their mean discharges, converted to head with compute_head's rule
(oneka/model.py:343-350) at the median conductivity and mean thickness, plus
N(0, 0.1 m) noise; every observation's standard deviation is 0.1 m.
Observations closer than 100 m to a well are redrawn (the reference's driver
drops them, oneka/utilities.py:34-144).
"""
import math

import numpy as np

from .probabilityfield import ProbabilityField
from .stochastic import Lognormal, Signed, compute_variate_mean

TRUE_COEF = (-1.0e-5, -1.0e-5, 0.0, 1.0e-3, 1.0e-3, 1000.0)
C4PI = 0.07957747154594767


class Site:
    """One capture-zone problem: inputs plus run sizes."""

    def __init__(self, name, wells, c_dist, p_dist, t_dist, observations, base, npaths, duration,
                 nrealizations, spacing, nrows, ncols, umbra, tol=1.0, maxstep=10.0,
                 confined=True):
        self.name = name
        self.wells = wells               # [(xw, yw, rw, q_dist)], target first
        self.target = 0
        self.c_dist, self.p_dist, self.t_dist = c_dist, p_dist, t_dist
        self.observations = observations
        self.base = base
        self.npaths = npaths
        self.duration = duration
        self.nrealizations = nrealizations
        self.spacing = spacing
        self.nrows, self.ncols = nrows, ncols
        self.umbra = umbra
        self.tol, self.maxstep = tol, maxstep
        self.confined = confined

    @property
    def xt(self):
        return self.wells[self.target][0]

    @property
    def yt(self):
        return self.wells[self.target][1]

    @property
    def rt(self):
        return self.wells[self.target][2]

    def wells_xy(self):
        return np.array([[w[0], w[1]] for w in self.wells], dtype=np.float64)

    def obs_array(self):
        return np.array([o[:4] for o in self.observations], dtype=np.float64)

    def fixed_field(self):
        """A ProbabilityField pre-expanded to nrows x ncols nodes around the
        target, built with the reference's own calls (constructor + expand)."""
        return fixed_field(self.xt, self.yt, self.spacing, self.nrows, self.ncols)


def fixed_bounds(xt, yt, spacing, nrows, ncols):
    """The (xmin, xmax, ymin, ymax) whose expand() of the 3 x 3 field around
    (xt, yt) yields nrows x ncols nodes, ncols // 2 of them left of the target."""
    left, below = ncols // 2, nrows // 2
    right, above = ncols - 1 - left, nrows - 1 - below
    return (xt - (left - 0.5) * spacing, xt + (right - 0.5) * spacing,
            yt - (below - 0.5) * spacing, yt + (above - 0.5) * spacing)


def fixed_field(xt, yt, spacing, nrows, ncols):
    pf = ProbabilityField(spacing, spacing, xt, yt)
    pf.expand(*fixed_bounds(xt, yt, spacing, nrows, ncols))
    assert (pf.nrows, pf.ncols) == (nrows, ncols)
    return pf


def _truth_heads(xy, wells_xyq, xt, yt, K, b, base):
    A, B, C, D, E, F = TRUE_COEF
    heads = []
    for x, y in xy:
        dx, dy = x - xt, y - yt
        phi = A * dx * dx + B * dy * dy + C * dx * dy + D * dx + E * dy + F
        for xw, yw, q in wells_xyq:
            phi += q * math.log((x - xw) ** 2 + (y - yw) ** 2) * C4PI
        if phi <= 0:
            raise ValueError("true potential is not positive at an observation")
        if phi < 0.5 * K * b * b:
            h = math.sqrt(2.0 * phi / K)
        else:
            h = (phi + 0.5 * K * b * b) / (K * b)
        heads.append(base + h)
    return np.array(heads)


def _place(rng, n, lo, hi, avoid, min_dist):
    pts = []
    while len(pts) < n:
        p = rng.uniform(lo, hi, 2)
        if all((p[0] - a[0]) ** 2 + (p[1] - a[1]) ** 2 >= min_dist ** 2 for a in avoid + pts):
            pts.append((float(p[0]), float(p[1])))
    return pts


def _make(name, seed, domain, nwells, q_dist_fn, nobs, npaths, years, nreal, spacing, nside,
          rw=0.25):
    rng = np.random.default_rng(seed)
    lo, hi = domain
    xt = yt = 0.5 * (lo + hi)
    others = _place(rng, nwells - 1, lo, hi, [(xt, yt)], 50.0)
    wells = [(xt, yt, rw, q_dist_fn(0))]
    wells += [(x, y, rw, q_dist_fn(i + 1)) for i, (x, y) in enumerate(others)]
    c_dist = Lognormal(math.log(10.0), 0.5)
    p_dist = (0.2, 0.3)
    t_dist = (20.0, 30.0)
    base = 0.0
    wxy = [(w[0], w[1]) for w in wells]
    obs_xy = _place(rng, nobs, lo, hi, wxy, 100.0)
    wq = [(w[0], w[1], compute_variate_mean(w[3])) for w in wells]
    heads = _truth_heads(obs_xy, wq, xt, yt, 10.0, compute_variate_mean(t_dist), base)
    heads = heads + rng.normal(0.0, 0.1, nobs)
    observations = [(x, y, float(z), 0.1) for (x, y), z in zip(obs_xy, heads)]
    return Site(name, wells, c_dist, p_dist, t_dist, observations, base, npaths,
                years * 365.25, nreal, spacing, nside, nside, 2.0 * spacing)


def config0():
    """4 km square, target + 3 wells, q ~ triangular(300, 500, 700), 10
    piezometers; 100 realizations x 100 particles, 10 years, 100^2 nodes at 40 m."""
    return _make("cfg0: 4 wells, 100 x 100 particles, 10 y, 100^2 @ 40 m", 20260000,
                 (0.0, 4000.0), 4, lambda i: (300.0, 500.0, 700.0), 10, 100, 10, 100, 40.0, 100)


def config1(nreal=10000):
    """8 km square, 25 wells, q ~ U(100, 1500), 40 piezometers; 10^4
    realizations x 1000 particles, 20 years, 400^2 nodes at 20 m."""
    return _make("cfg1: 25 wells, 10^4 x 1000 particles, 20 y, 400^2 @ 20 m", 20260001,
                 (0.0, 8000.0), 25, lambda i: (100.0, 1500.0), 40, 1000, 20, nreal, 20.0, 400)


def config4(nreal=1000000):
    """20 km square, 200 wells, q ~ lognormal(ln 800, 0.6) with 10 % injection,
    200 piezometers; 10^6 realizations x 256 particles, 30 years, 2000^2 nodes
    at 10 m."""
    return _make("cfg4: 200 wells, 10^6 x 256 particles, 30 y, 2000^2 @ 10 m", 20260004,
                 (0.0, 20000.0), 200,
                 lambda i: Lognormal(math.log(800.0), 0.6) if i == 0 else
                 Signed(Lognormal(math.log(800.0), 0.6), 0.1),
                 200, 256, 30, nreal, 10.0, 2000)


def config2(nsites=64, scale=1.0, seed=20260002):
    """64 independent sites: wells U{1..200} (the target first; q ~ U(100,
    1500)), observations U{10..100}, a square lattice of U{100..800} nodes a
    side at 20 m, realizations U{100..5000} (times `scale`, at least 1) x 256
    particles, backtraces of U(1, 50) years.  Each site's domain grows with its
    well count (8 km for 25 wells, at least 4 km) so that wells stay 50 m and
    observations 100 m apart."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nsites):
        nw = int(rng.integers(1, 201))
        nobs = int(rng.integers(10, 101))
        nside = int(rng.integers(100, 801))
        nreal = max(1, int(round(int(rng.integers(100, 5001)) * scale)))
        years = float(rng.uniform(1.0, 50.0))
        side = max(4000.0, 8000.0 * math.sqrt(nw / 25.0))
        out.append(_make(f"cfg2 site {k}: {nw} wells, {nobs} obs, {nreal} x 256, {years:.1f} y, "
                         f"{nside}^2 @ 20 m", seed * 100 + k, (0.0, side), nw,
                         lambda i: (100.0, 1500.0), nobs, 256, years, nreal, 20.0, nside))
    return out


def draw(site, nreal=None, seed=0):
    """Seeded realization parameters for a site: dict(q, K, n, b, z)."""
    from .stochastic import draw_realizations
    rng = np.random.default_rng(seed)
    return draw_realizations(site.wells, site.c_dist, site.p_dist, site.t_dist,
                             site.nrealizations if nreal is None else nreal, rng)
