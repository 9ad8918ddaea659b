"""Oneka-type analytic element model (oneka/model.py:62-606), evaluated on the GPU.

``Model`` keeps the reference's constructor, attributes and methods.  Point
evaluations (``compute_potential`` ... ``compute_velocity_confined``) and the
fit (``fit_regional_flow``) run in liboka's FP64 kernels in the reference's
operation order; they also accept arrays of points.  ``construct_fit`` and
``compute_fit`` are kept as host helpers for API compatibility; the hot path
fits through ``oka_fit_draw`` instead.
"""
import logging

import numpy as np

from . import _lib

log = logging.getLogger("Oneka")


class Error(Exception):
    """Base class for module errors (model.py:46-48)."""


class RangeError(Error):
    """Passed argument out of valid range (model.py:51-53)."""


class AquiferError(Error):
    """The aquifer is dry: head <= 0 (model.py:56-58)."""


class Model:
    """An Oneka-type analytic element model (model.py:62-197).

    wells is a list of (xw, yw, rw, qw); (xo, yo) is the local origin of the
    regional flow; coef holds A..F of the regional potential
    A dx^2 + B dy^2 + C dx dy + D dx + E dy + F.
    """

    def __init__(self, base, conductivity, porosity, thickness, wells, xo=0, yo=0,
                 coef=np.zeros((6,))):
        self.base = base
        self.conductivity = conductivity
        self.porosity = porosity
        self.thickness = thickness
        self.wells = wells
        self.xo = xo
        self.yo = yo
        self.coef = coef

    @property
    def coef(self):
        return self._coef

    @coef.setter
    def coef(self, coef):
        self._coef = np.reshape(coef, [6, ])

    def __repr__(self):
        return 'Model({0.base}, {0.conductivity}, {0.porosity}, {0.thickness})'.format(self)

    __str__ = __repr__

    # -- GPU evaluation -----------------------------------------------------
    def _wells_array(self):
        w = np.zeros((len(self.wells), 4))
        for i, well in enumerate(self.wells):
            w[i] = [float(well[0]), float(well[1]), float(well[2]), float(well[3])]
        return w

    def _struct(self, wells):
        m = _lib.Model_()
        m.base = float(self.base)
        m.conductivity = float(self.conductivity)
        m.porosity = float(self.porosity)
        m.thickness = float(self.thickness)
        m.xo, m.yo = float(self.xo), float(self.yo)
        for j in range(6):
            m.coef[j] = float(self.coef[j])
        m.nwells = wells.shape[0]
        m.wells = _lib.ptr(wells)
        return m

    def _eval(self, kind, x, y):
        scalar = np.ndim(x) == 0 and np.ndim(y) == 0
        xb, yb = np.broadcast_arrays(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
        xy = np.ascontiguousarray(np.stack([xb.ravel(), yb.ravel()], axis=1))
        n = xy.shape[0]
        width = 2 if kind in (_lib.EVAL_DISCHARGE, _lib.EVAL_VELOCITY,
                              _lib.EVAL_VELOCITY_CONFINED) else 1
        out = np.zeros((n, width))
        status = np.zeros(n, dtype=np.int32)
        wells = self._wells_array()
        m = self._struct(wells)
        ctx = _lib.context()
        ctx.check(ctx.lib.oka_model_eval(ctx.h, m, kind, n, _lib.ptr(xy), _lib.ptr(out),
                                         _lib.ptr(status), 0))
        if np.any(status == _lib.ERR_AQUIFER):
            if kind == _lib.EVAL_HEAD:
                raise AquiferError("potential_to_head: potential <= 0")
            raise AquiferError("discharge_to_velocity: head <= 0")
        if scalar:
            return out[0, 0] if width == 1 else (out[0, 0], out[0, 1])
        shape = xb.shape
        if width == 1:
            return out[:, 0].reshape(shape)
        return out[:, 0].reshape(shape), out[:, 1].reshape(shape)

    def compute_potential(self, x, y):
        """Discharge potential [m^3/d] at (x, y) (model.py:207-237)."""
        return self._eval(_lib.EVAL_POTENTIAL, x, y)

    def compute_potential_wells_only(self, x, y):
        """Potential of the wells alone (model.py:240-266)."""
        return self._eval(_lib.EVAL_POTENTIAL_WELLS, x, y)

    def compute_discharge(self, x, y):
        """[Qx, Qy], vertically integrated discharge [m^2/d] (model.py:269-315)."""
        qx, qy = self._eval(_lib.EVAL_DISCHARGE, x, y)
        return [qx, qy]

    def compute_head(self, x, y):
        """Piezometric head above the base [m] (model.py:318-350)."""
        return self._eval(_lib.EVAL_HEAD, x, y)

    def compute_velocity(self, x, y):
        """(Vx, Vy), vertically averaged seepage velocity [m/d] (model.py:353-389)."""
        return self._eval(_lib.EVAL_VELOCITY, x, y)

    def compute_velocity_confined(self, x, y):
        """(Vx, Vy) assuming a confined aquifer (model.py:392-427)."""
        return self._eval(_lib.EVAL_VELOCITY_CONFINED, x, y)

    # -- fit ----------------------------------------------------------------
    def fit_regional_flow(self, obs, xo, yo):
        """Weighted least-squares fit of A..F (model.py:430-494).

        Returns (coef_ev (6, 1), coef_cov (6, 6)) and sets xo, yo, coef.
        Computed by liboka's FP64 fit kernel (normal equations with column
        equilibration, Cholesky, one refinement step; cov = inv(WA^T WA)).
        """
        obs = _lib.f64([[float(v) for v in o[:4]] for o in obs]).reshape(-1, 4)
        wells = self._wells_array()
        wxy = np.ascontiguousarray(wells[:, :2])
        q = np.ascontiguousarray(wells[:, 3])
        site = _lib.make_site(wxy, xo, yo, self.base, obs)
        K = np.array([float(self.conductivity)])
        b = np.array([float(self.thickness)])
        ev = np.zeros(6)
        cov = np.zeros(36)
        coef = np.zeros(6)
        status = np.zeros(1, dtype=np.int32)
        ctx = _lib.context()
        rc = ctx.lib.oka_fit_draw(ctx.h, site, 1, _lib.ptr(K), _lib.ptr(b), _lib.ptr(q), None,
                                  _lib.ptr(ev), _lib.ptr(cov), _lib.ptr(coef), _lib.ptr(status), 0)
        if rc == _lib.ERR_RANGE:
            raise RangeError("model.fit_coeficient: elevation < base")
        if rc == _lib.ERR_LINALG:
            log.error(' numpy.linalg.lstsq: failed')
            raise np.linalg.LinAlgError(ctx.lib.oka_last_error(ctx.h).decode())
        ctx.check(rc)
        coef_ev = ev.reshape(6, 1)
        coef_cov = cov.reshape(6, 6)
        self.xo = xo
        self.yo = yo
        self.coef = coef_ev
        return (coef_ev, coef_cov)

    def construct_fit(self, obs, xo, yo):
        """(WA, Wb) of the weighted least squares (model.py:497-567), on the host."""
        nobs = len(obs)
        WA = np.zeros([nobs, 6])
        Wb = np.zeros([nobs, 1])
        xs = np.array([float(o[0]) for o in obs])
        ys = np.array([float(o[1]) for o in obs])
        pw = self.compute_potential_wells_only(xs, ys) if nobs else np.zeros(0)
        K, b = self.conductivity, self.thickness
        for i in range(nobs):
            x, y, z_ev, z_std = obs[i][:4]
            head = z_ev - self.base
            if head >= b:
                pot_ev = K * b * (head - 0.5 * b)
                pot_std = K * b * z_std
            elif head > 0:
                pot_ev = 0.5 * K * (head * head + z_std * z_std)
                pot_std = K * head * z_std
            else:
                raise RangeError("model.fit_coeficient: elevation < base")
            dx = x - xo
            dy = y - yo
            WA[i, :] = np.array([dx * dx, dy * dy, dx * dy, dx, dy, 1]) / pot_std
            Wb[i] = (pot_ev - pw[i]) / pot_std
        return (WA, Wb)

    @staticmethod
    def compute_fit(WA, Wb):
        """(coef_ev, coef_cov) from (WA, Wb) (model.py:570-606), on the host."""
        try:
            coef_ev = np.linalg.lstsq(WA, Wb, rcond=-1)[0]
        except np.linalg.LinAlgError:
            log.error(' numpy.linalg.lstsq: failed')
            raise
        coef_cov = np.linalg.inv(np.matmul(WA.T, WA))
        return (coef_ev, coef_cov)

    # -- drop-in feval --------------------------------------------------------
    def velocity_function(self, confined=True):
        """The backtracing feval of stochastic.py:253-260 bound to this model."""
        from .capturezone import ModelVelocity
        return ModelVelocity(self, confined)
