"""Single-well capture zone (oneka/capturezone.py:51-253) on the GPU.

``compute_capturezone`` and ``compute_backtrace`` keep the reference's
signatures.  The velocity function ``feval`` must be a ``ModelVelocity`` (what
``Model.velocity_function(confined)`` returns): it is callable like the
reference's closures (stochastic.py:253-260), and it tells the GPU which model
to trace through.  An arbitrary Python callable cannot run on the GPU and is
rejected with a TypeError.
"""
import logging

import numpy as np

from . import _lib
from .engine import Realizations, chronicle_into_field, warn_statuses

log = logging.getLogger(__name__)

STEPAWAY = 1.0   # distance beyond the well radius [m] (capturezone.py:110)


class ModelVelocity:
    """feval(xy) = -V(xy) of a Model (stochastic.py:253-260).

    confined=True uses compute_velocity_confined, False compute_velocity.
    """

    def __init__(self, model, confined=True):
        self.model = model
        self.confined = bool(confined)

    def __call__(self, xy):
        m = self.model
        f = m.compute_velocity_confined if self.confined else m.compute_velocity
        vx, vy = f(xy[0], xy[1])
        return np.array([-vx, -vy])

    def site_arrays(self):
        """(wells_xy [nw, 2], Realizations of this single model)."""
        m = self.model
        w = np.array([[float(v) for v in well[:4]] for well in m.wells],
                     dtype=np.float64).reshape(-1, 4)
        real = Realizations([float(m.conductivity)], [float(m.porosity)], [float(m.thickness)],
                            w[:, 3].reshape(1, -1), np.asarray(m.coef, dtype=np.float64))
        return np.ascontiguousarray(w[:, :2]), real


def _model_velocity(feval):
    if not isinstance(feval, ModelVelocity):
        raise TypeError("feval must be a ModelVelocity (Model.velocity_function(confined)); "
                        "an arbitrary Python callable cannot be traced on the GPU")
    return feval


def start_points(xtarget, ytarget, rtarget, npaths):
    """The npaths start points around the target (capturezone.py:109-115), one
    numpy scalar evaluation per angle, as the reference computes them."""
    pts = np.zeros((npaths, 2))
    for k, theta in enumerate(np.linspace(0, 2 * np.pi, npaths + 1)[0:-1]):
        pts[k, 0] = (rtarget + STEPAWAY) * np.cos(theta) + xtarget
        pts[k, 1] = (rtarget + STEPAWAY) * np.sin(theta) + ytarget
    return pts


def compute_capturezone(xtarget, ytarget, rtarget, npaths, duration, pfield, umbra, weight, tol,
                        maxstep, feval, precision="fp64", max_steps=0):
    """Backtrack npaths particles from around the target and register their
    traces in pfield with the given weight (capturezone.py:51-123).

    precision "fp64" is the FP64-exact tracer; "fp32" the FP32 tracer.
    """
    fv = _model_velocity(feval)
    wells_xy, real = fv.site_arrays()
    starts = start_points(xtarget, ytarget, rtarget, npaths)
    res = chronicle_into_field(pfield, wells_xy, float(fv.model.xo), float(fv.model.yo), real,
                               starts, duration, tol, maxstep, umbra, fv.confined,
                               _precision(precision), max_steps)
    warn_statuses(res.trace_status)
    if not pfield.empty:
        pfield.rgrid |= res.counts > 0
    pfield.register(weight)


def compute_backtrace(xs, ys, duration, tol, maxstep, feval, max_steps=0):
    """Vertices of one Dormand-Prince backtrace from (xs, ys) (capturezone.py:127-253)."""
    fv = _model_velocity(feval)
    wells_xy, real = fv.site_arrays()
    verts, nverts, status = trace_vertices(wells_xy, float(fv.model.xo), float(fv.model.yo), real,
                                           np.array([[float(xs), float(ys)]]), duration, tol,
                                           maxstep, fv.confined, max_steps)
    if status[0] != _lib.TRACE_OK:
        warn_statuses(status)
    return [(float(x), float(y)) for x, y in verts[0][:nverts[0]]]


def trace_vertices(wells_xy, xo, yo, real, starts, duration, tol, maxstep, confined=True,
                   max_steps=0, cap=4096):
    """All vertices of every (realization, start point) trace, FP64-exact.

    Returns (verts [R*P, cap, 2], nverts [R*P], status [R*P]); cap grows to the
    longest trace.
    """
    wells_xy = _lib.f64(wells_xy).reshape(-1, 2)
    starts = _lib.f64(starts).reshape(-1, 2)
    ntr = len(real) * starts.shape[0]
    ctx = _lib.context()
    site = _lib.make_site(wells_xy, xo, yo)
    opts = _lib.make_trace_opts(starts.shape[0], duration, tol, maxstep, 0.0, _lib.ptr(starts),
                                _lib.FP64_EXACT, confined, max_steps)
    while True:
        verts = np.zeros((ntr, cap, 2))
        nverts = np.zeros(ntr, dtype=np.int64)
        status = np.zeros(ntr, dtype=np.int32)
        ctx.check(ctx.lib.oka_trace_vertices(
            ctx.h, site, len(real), _lib.ptr(real.K), _lib.ptr(real.n), _lib.ptr(real.b),
            _lib.ptr(real.q), _lib.ptr(real.coef), opts, cap, _lib.ptr(verts), _lib.ptr(nverts),
            _lib.ptr(status), None, 0))
        if ntr == 0 or nverts.max() <= cap:
            return verts, nverts, status
        cap = int(nverts.max())


def _precision(p):
    if p in ("fp64", _lib.FP64_EXACT):
        return _lib.FP64_EXACT
    if p in ("fp32", _lib.FP32):
        return _lib.FP32
    raise ValueError("precision must be 'fp64' or 'fp32'")
