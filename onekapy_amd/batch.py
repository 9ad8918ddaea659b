"""Many sites in one persistent, graph-captured GPU run (oka_batch_*).

Each site is a ``create_stochastic_capturezone`` problem (oneka/stochastic.py:
76-270) on a fixed, pre-expanded ProbabilityField, tracing its own velocity
field (``site.confined``).  The library stages every site once, captures one
run -- per site the fit and MVN draw of every realization, then all particles
of all sites in one launch of the pool tracer (work stealing over live
particles, FP32 or FP64-exact for the whole batch) -- as a HIP graph, and
replays it.
"""
import ctypes as C

import numpy as np

from . import _lib
from .capturezone import start_points

KNOBS = ("pool_slots", "pool_waves_per_cu", "pool_chunk")


def _knob_opts(knobs):
    o = _lib.TraceOpts_()
    for k in KNOBS:
        setattr(o, k, int((knobs or {}).get(k, 0)))
    return o


class SiteBatch:
    """sites: objects with the attributes of ``sites.Site``; draws: per site a
    dict(q, K, n, b, z) of its realization parameters; precision: the batch's
    tracer arithmetic (_lib.FP32 or _lib.FP64_EXACT); far_order: every site's
    oka_trace_opts.far_order (FP32, confined sites: distant wells as local
    expansions; 0 lets the library choose, < 0 sums every well directly)."""

    def __init__(self, sites, draws, knobs=None, device=0, precision=_lib.FP32, far_order=0):
        self.ctx = _lib.context(device)
        self.sites = list(sites)
        self.knobs = dict(knobs or {})
        keep = []
        arr = (_lib.BatchSite_ * len(self.sites))()
        for i, (s, d) in enumerate(zip(self.sites, draws)):
            wxy = np.ascontiguousarray(s.wells_xy(), dtype=np.float64)
            obs = np.ascontiguousarray(s.obs_array(), dtype=np.float64)
            st = _lib.make_site(wxy, s.xt, s.yt, s.base, obs)
            starts = np.ascontiguousarray(start_points(s.xt, s.yt, s.rt, s.npaths))
            arrays = [np.ascontiguousarray(d[k], dtype=np.float64) for k in ("K", "n", "b", "q")]
            z = None if d.get("z") is None else np.ascontiguousarray(d["z"], dtype=np.float64)
            pf = s.fixed_field()
            b = arr[i]
            b.site = C.pointer(st)
            b.nreal = len(arrays[0])
            b.K, b.n, b.b, b.q = (_lib.ptr(a) for a in arrays)
            b.z = _lib.ptr(z)
            b.opts = _lib.make_trace_opts(s.npaths, s.duration, s.tol, s.maxstep, s.umbra,
                                          _lib.ptr(starts), precision, s.confined,
                                          far_order=far_order)
            b.grid = _lib.make_grid(pf.xmin, pf.ymin, pf.xmax, pf.ymax, pf.deltax, pf.deltay,
                                    pf.nrows, pf.ncols)
            keep.append((wxy, obs, st, starts, arrays, z))
        h = C.c_void_p()
        opts = _knob_opts(self.knobs)
        self.ctx.check(self.ctx.lib.oka_batch_create(self.ctx.h, len(self.sites), C.addressof(arr),
                                                     C.byref(opts), C.byref(h)))
        self.h = h
        self.nreal = [int(arr[i].nreal) for i in range(len(self.sites))]

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.oka_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, **knobs):
        """Re-captures the run with other pool knobs (pool_slots,
        pool_waves_per_cu, pool_chunk; 0 = the library's default)."""
        self.knobs = {k: int(knobs.get(k, 0)) for k in KNOBS}
        opts = _knob_opts(self.knobs)
        self.ctx.check(self.ctx.lib.oka_batch_tune(self.h, C.byref(opts)))

    def run(self):
        """One replay of the captured run; returns its device time in ms."""
        ms = C.c_float(0.0)
        self.ctx.check(self.ctx.lib.oka_batch_run(self.h, C.byref(ms)))
        return float(ms.value)

    def results(self, s):
        """(counts [nrows, ncols] int32, coef [nreal, 6], TraceStats_) of site s."""
        site = self.sites[s]
        counts = np.zeros((site.nrows, site.ncols), dtype=np.int32)
        coef = np.zeros((self.nreal[s], 6))
        st = _lib.TraceStats_()
        self.ctx.check(self.ctx.lib.oka_batch_results(self.h, s, counts.ctypes.data,
                                                      coef.ctypes.data, C.byref(st)))
        return counts, coef, st

    def fields(self):
        """Each site's ProbabilityField after the last run: pgrid holds the
        number of realizations that marked each node, total_weight = R
        (stochastic.py:262-270 with weight 1)."""
        out = []
        for s, site in enumerate(self.sites):
            counts, _, _ = self.results(s)
            pf = site.fixed_field()
            pf.pgrid += counts.astype(np.float64)
            pf.total_weight += float(self.nreal[s])
            out.append(pf)
        return out
