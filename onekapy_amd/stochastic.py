"""Monte Carlo capture zone (oneka/stochastic.py:76-376) on the GPU.

``create_stochastic_capturezone`` keeps the reference's signature and its six
steps per realization (stochastic.py:187-207); the work is batched over all
realizations instead of looping:

1. all realization parameters are drawn on the host from a seeded generator
   (discharges, conductivity, porosity, thickness, and the standard normals of
   the MVN draw);
2. one GPU launch fits every realization's regional flow and draws its A..F
   (``oka_fit_draw``);
3. the engine traces every particle of every realization and rasterizes the
   traces with the reference's expansion history (``engine.py``).

The distribution vocabulary of generate_random_variate (stochastic.py:274-306)
is kept -- scalar = constant, (lo, hi) = uniform, (lo, mode, hi) = triangular
-- and extended with ``Lognormal`` and ``Signed`` for the benchmark sites.
The reference draws from numpy's unseeded global RandomState and an unseeded
default_rng per realization; here one seeded numpy Generator makes a run
reproducible.
"""
import ctypes as C
import logging

import numpy as np

from . import _lib
from .capturezone import start_points, _precision
from .engine import Realizations, chronicle_into_field, warn_statuses
from .probabilityfield import ProbabilityField

log = logging.getLogger("Oneka")


class Error(Exception):
    """Base class for module errors (stochastic.py:65-67)."""


class DistributionError(Error):
    """Invalid distribution (stochastic.py:70-72)."""


class Lognormal:
    """exp(N(mu, sigma^2)); mu and sigma of the underlying normal."""

    def __init__(self, mu, sigma):
        self.mu, self.sigma = float(mu), float(sigma)

    def sample(self, rng, size=None):
        return rng.lognormal(self.mu, self.sigma, size)

    def mean(self):
        return float(np.exp(self.mu + 0.5 * self.sigma ** 2))

    def bounds(self):
        return (0.0, np.inf)


class Signed:
    """A distribution whose sign is flipped with probability p_negative
    (injection wells among discharge wells)."""

    def __init__(self, dist, p_negative):
        self.dist, self.p = dist, float(p_negative)

    def sample(self, rng, size=None):
        v = _sample(self.dist, rng, size)
        flip = rng.random(size) < self.p
        return np.where(flip, -v, v) if size is not None else (-v if flip else v)

    def mean(self):
        return (1.0 - 2.0 * self.p) * compute_variate_mean(self.dist)

    def bounds(self):
        return (-np.inf, np.inf)


def _sample(arg, rng, size=None):
    if hasattr(arg, "sample"):
        return arg.sample(rng, size)
    if type(arg) is not tuple:
        return arg if size is None else np.full(size, arg, dtype=np.float64)
    if len(arg) == 2:
        return rng.uniform(arg[0], arg[1], size)
    if len(arg) == 3:
        return rng.triangular(arg[0], arg[1], arg[2], size)
    raise DistributionError('<arg> must be a scalar, pair, or triple.')


def generate_random_variate(arg, rng=None):
    """A variate of a constant, uniform or triangular distribution (stochastic.py:274-306)."""
    if rng is None:
        rng = np.random.default_rng()
    return _sample(arg, rng)


def compute_variate_mean(arg):
    """The mean of a constant, uniform or triangular distribution (stochastic.py:310-342)."""
    if hasattr(arg, "mean"):
        return arg.mean()
    if type(arg) is not tuple:
        return arg
    if len(arg) == 2:
        return (arg[0] + arg[1]) / 2.0
    if len(arg) == 3:
        return (arg[0] + arg[1] + arg[2]) / 3.0
    raise DistributionError('<arg> must be a scalar, pair, or triple.')


def isdistribution(arg, lb, ub):
    """Do the arguments define a valid distribution within [lb, ub]? (stochastic.py:345-376)"""
    if hasattr(arg, "bounds"):
        lo, hi = arg.bounds()
        return lb <= lo and hi <= ub
    if isinstance(arg, int) or isinstance(arg, float):
        return lb <= arg <= ub
    if isinstance(arg, tuple) or isinstance(arg, list):
        if len(arg) == 2:
            return lb <= arg[0] <= arg[1] <= ub
        elif len(arg) == 3:
            return lb <= arg[0] <= arg[1] <= arg[2] <= ub
    return False


def oka_dist(arg):
    """The library's description (oka_dist) of one distribution of the vocabulary."""
    d = _lib.Dist_()
    if isinstance(arg, Signed):
        d = oka_dist(arg.dist)
        d.p_negative = arg.p
        return d
    if isinstance(arg, Lognormal):
        d.kind, d.a, d.b = _lib.DIST_LOGNORMAL, arg.mu, arg.sigma
    elif type(arg) is not tuple:
        d.kind, d.a = _lib.DIST_CONST, float(arg)
    elif len(arg) == 2:
        d.kind, d.a, d.b = _lib.DIST_UNIFORM, float(arg[0]), float(arg[1])
    elif len(arg) == 3:
        d.kind, d.a, d.b, d.c = _lib.DIST_TRIANGULAR, float(arg[0]), float(arg[1]), float(arg[2])
    else:
        raise DistributionError('<arg> must be a scalar, pair, or triple.')
    return d


def oka_dists(stochastic_wells, c_dist, p_dist, t_dist):
    """Array of oka_dist: each well's discharge, then K, porosity, thickness."""
    items = [oka_dist(w[3]) for w in stochastic_wells] + [oka_dist(c_dist), oka_dist(p_dist),
                                                          oka_dist(t_dist)]
    return (_lib.Dist_ * len(items))(*items)


def draw_realizations_device(stochastic_wells, c_dist, p_dist, t_dist, nrealizations, seed,
                             first=0, antithetic=False, qmc=False):
    """The same draws as draw_realizations, from the library's Philox4x32-10
    stream `seed` (realizations first .. first + n - 1).  With antithetic
    (even count and first), realization 2k + 1 reflects 2k's draws; with qmc,
    realization j is Sobol' point j under the nested scramble keyed by `seed`."""
    R, nw = int(nrealizations), len(stochastic_wells)
    out = dict(q=np.zeros((R, nw)), K=np.zeros(R), n=np.zeros(R), b=np.zeros(R),
               z=np.zeros((R, 6)))
    dists = oka_dists(stochastic_wells, c_dist, p_dist, t_dist)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_draw_realizations(
        ctx.h, nw, C.addressof(dists), R, int(seed) & (2 ** 64 - 1), int(first),
        _lib.ptr(out["q"]), _lib.ptr(out["K"]), _lib.ptr(out["n"]), _lib.ptr(out["b"]),
        _lib.ptr(out["z"]), (_lib.ANTITHETIC if antithetic else 0) | (_lib.QMC if qmc else 0)))
    return out


def draw_realizations(stochastic_wells, c_dist, p_dist, t_dist, nrealizations, rng):
    """Host draws of every realization's random inputs (stochastic.py:222-241).

    Returns dict(q [R, nw], K [R], n [R], b [R], z [R, 6] standard normals).
    The order is fixed -- discharges well by well, then K, n, b, then z -- so a
    seed reproduces a run.
    """
    R = int(nrealizations)
    nw = len(stochastic_wells)
    q = np.zeros((R, nw))
    for j, w in enumerate(stochastic_wells):
        q[:, j] = _sample(w[3], rng, R)
    K = np.asarray(_sample(c_dist, rng, R), dtype=np.float64)
    n = np.asarray(_sample(p_dist, rng, R), dtype=np.float64)
    b = np.asarray(_sample(t_dist, rng, R), dtype=np.float64)
    z = rng.standard_normal((R, 6))
    return dict(q=q, K=K, n=n, b=b, z=z)


def fit_and_draw(wells_xy, xo, yo, base, observations, K, b, q, z, moments=True):
    """GPU fit of every realization and its MVN draw (model.py:430-606, stochastic.py:239-241).

    Returns (ev [R, 6], cov [R, 6, 6], coef [R, 6]).  z may be None (coef = ev).
    With moments=False only coef is computed and ev, cov are None.
    """
    from .model import RangeError
    obs = _lib.f64([[float(v) for v in o[:4]] for o in observations]).reshape(-1, 4)
    wells_xy = _lib.f64(wells_xy).reshape(-1, 2)
    K, b = _lib.f64(K).reshape(-1), _lib.f64(b).reshape(-1)
    R = K.shape[0]
    q = _lib.f64(q).reshape(R, wells_xy.shape[0])
    zz = None if z is None else _lib.f64(z).reshape(R, 6)
    ev = np.zeros((R, 6)) if moments else None
    cov = np.zeros((R, 6, 6)) if moments else None
    coef = np.zeros((R, 6))
    status = np.zeros(R, dtype=np.int32)
    ctx = _lib.context()
    site = _lib.make_site(wells_xy, xo, yo, base, obs)
    rc = ctx.lib.oka_fit_draw(ctx.h, site, R, _lib.ptr(K), _lib.ptr(b), _lib.ptr(q), _lib.ptr(zz),
                              _lib.ptr(ev), _lib.ptr(cov), _lib.ptr(coef), _lib.ptr(status), 0)
    if rc == _lib.ERR_RANGE:
        raise RangeError("model.fit_coeficient: elevation < base")
    if rc == _lib.ERR_LINALG:
        log.error(' numpy.linalg.lstsq: failed')
        raise np.linalg.LinAlgError(ctx.lib.oka_last_error(ctx.h).decode())
    ctx.check(rc)
    return ev, cov, coef


def create_stochastic_capturezone(target, npaths, duration, nrealizations, base, c_dist, p_dist,
                                  t_dist, stochastic_wells, observations, spacing, umbra,
                                  confined, tol, maxstep, seed=None, precision="fp64",
                                  max_steps=0):
    """Compute the stochastic capture zone for the target well (stochastic.py:76-270).

    Returns the ProbabilityField, with pgrid holding, per node, the number of
    realizations whose traces came within umbra (weight 1.0 each) and
    total_weight = nrealizations.
    """
    xtarget, ytarget, rtarget = stochastic_wells[target][0:3]
    pfield = ProbabilityField(spacing, spacing, xtarget, ytarget)
    rng = np.random.default_rng(seed)
    d = draw_realizations(stochastic_wells, c_dist, p_dist, t_dist, nrealizations, rng)
    wells_xy = np.array([[float(w[0]), float(w[1])] for w in stochastic_wells]).reshape(-1, 2)
    ev, _cov, coef = fit_and_draw(wells_xy, xtarget, ytarget, base, observations, d["K"], d["b"],
                                  d["q"], d["z"])
    if log.isEnabledFor(logging.INFO):
        for i in range(int(nrealizations)):
            recharge = 2 * (coef[i, 0] + coef[i, 1])
            log.info('Realization #{0:d}: {1:.2f}, {2:.2f}, {3:.2f}, {4:.2f}, {5:.4e}'.format(
                i, base, d["K"][i], d["n"][i], d["b"][i], recharge))
            log.info('    coef ev:  ({0:+.4e}, {1:+.4e}, {2:+.4e}, {3:+.4e}, {4:+.4e}, {5:+.4e})'
                     .format(*ev[i]))
            log.info('    coef rng: ({0:+.4e}, {1:+.4e}, {2:+.4e}, {3:+.4e}, {4:+.4e}, {5:+.4e})'
                     .format(*coef[i]))
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(xtarget, ytarget, rtarget, npaths)
    res = chronicle_into_field(pfield, wells_xy, float(xtarget), float(ytarget), real, starts,
                               duration, tol, maxstep, umbra, confined, _precision(precision),
                               max_steps)
    warn_statuses(res.trace_status)
    R = int(nrealizations)
    pfield.pgrid += res.counts.astype(np.float64)
    pfield.total_weight += float(R)
    return pfield
