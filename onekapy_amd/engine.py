"""Batched trace + chronicle with the reference's expansion-history semantics.

In the reference every trace is followed by ``pfield.rasterize`` which first
grows the grid to the trace's bounding box (probabilityfield.py:334-335) and
then marks nodes only inside the grid as it stands at that moment; traces are
processed realization by realization, particle by particle
(stochastic.py:220-265, capturezone.py:113-123).  This module reproduces that
exactly for a whole batch:

1. one fused trace + rasterize launch on the field's current grid; if no trace
   would have grown the grid (its bbox strictly inside), the result is exact;
2. otherwise the per-trace bboxes of that launch go through the library's
   expansion-history scan (``oka_grid_history``), the field adopts the grown
   geometry, and a second launch re-traces with each trace clipped to the grid
   as it stood when that trace was rasterized.

All tracing and rasterization run on the GPU through liboka.
"""
import logging

import numpy as np

from . import _lib

log = logging.getLogger(__name__)

STATUS_TEXT = {
    _lib.TRACE_TRUNCATED: "trace terminated prematurely",
    _lib.TRACE_STALLED: "trace stalled (step size became NaN or zero)",
    _lib.TRACE_STEP_CAP: "trace stopped at the attempted-step cap",
}


class Realizations:
    """Per-realization model parameters of one site (host arrays)."""

    def __init__(self, K, n, b, q, coef):
        self.K = _lib.f64(K).reshape(-1)
        R = self.K.shape[0]
        self.n = _lib.f64(n).reshape(R)
        self.b = _lib.f64(b).reshape(R)
        q = _lib.f64(q)
        nw = q.shape[-1] if q.ndim == 2 else (q.size // R if R else 0)
        self.q = q.reshape(R, nw)
        self.coef = _lib.f64(coef).reshape(R, 6)

    def __len__(self):
        return self.K.shape[0]


class ChronicleResult:
    def __init__(self, counts, status, stats, passes, bbox=None):
        self.counts = counts          # int32 [nrows, ncols] on the field's final grid
        self.trace_status = status    # int32 [R * P]
        self.bbox = bbox              # float64 [R * P, 4]: xmin, xmax, ymin, ymax per trace
        self.stats = stats            # _lib.TraceStats_ of the rasterizing launch
        self.passes = passes          # 1 (fused pass was exact) or 2 (history re-trace)


def _launch(ctx, site, real, opts, grid, epochs, rasterize, counts, bbox, status, stats):
    ne = 0 if epochs is None else len(epochs)
    eptr = None if epochs is None else epochs.ctypes.data
    ctx.check(ctx.lib.oka_trace_chronicle(
        ctx.h, site, len(real), _lib.ptr(real.K), _lib.ptr(real.n), _lib.ptr(real.b),
        _lib.ptr(real.q), _lib.ptr(real.coef), opts, grid, ne, eptr, 1 if rasterize else 0,
        _lib.ptr(counts), _lib.ptr(bbox), _lib.ptr(status), None, stats, 0))


def chronicle_into_field(pfield, wells_xy, xo, yo, real, starts, duration, tol, maxstep, umbra,
                         confined=True, precision=_lib.FP64_EXACT, max_steps=0,
                         bitmap_mode=_lib.BITMAP_AUTO, tile_slots=0, schedule=_lib.SCHED_AUTO,
                         pool_slots=0, pool_waves_per_cu=0, pool_chunk=0, far_order=0):
    """Trace all realizations and rasterize them into ``pfield``'s lattice.

    Grows ``pfield`` exactly as the reference's per-trace ``rasterize`` calls
    would, and returns the per-node number of realizations that marked each
    node (the caller registers them).  Nothing is registered here.
    bitmap_mode / tile_slots place the per-realization visited-node bitmap
    and schedule / pool_slots / pool_waves_per_cu / pool_chunk choose how particles are
    scheduled (oka_trace_opts); every choice gives the same counts when every
    well is summed directly.  far_order (FP32 pool, confined) sums distant
    wells as local expansions: 0 lets the library choose (on from 64 wells,
    and then the automatic schedule is the pool), < 0 sums every well
    directly, as the workgroup schedule does.
    """
    wells_xy = _lib.f64(wells_xy).reshape(-1, 2)
    starts = _lib.f64(starts).reshape(-1, 2)
    R, P = len(real), starts.shape[0]
    ntr = R * P
    ctx = _lib.context()
    site = _lib.make_site(wells_xy, xo, yo)
    opts = _lib.make_trace_opts(P, duration, tol, maxstep, umbra, _lib.ptr(starts), precision,
                                confined, max_steps, bitmap_mode, tile_slots, schedule,
                                pool_slots, pool_waves_per_cu, pool_chunk, far_order)
    bbox = np.zeros((ntr, 4))
    status = np.zeros(ntr, dtype=np.int32)
    stats = _lib.TraceStats_()
    if ntr == 0:
        return ChronicleResult(np.zeros((pfield.nrows, pfield.ncols), np.int32), status, stats, 0,
                               bbox)
    if not pfield.empty:
        g = pfield.grid()
        counts = np.zeros((pfield.nrows, pfield.ncols), dtype=np.int32)
        _launch(ctx, site, real, opts, g, None, True, counts, bbox, status, stats)
        if stats.escaped == 0:
            return ChronicleResult(counts, status, stats, 1, bbox)
    else:
        _launch(ctx, site, real, opts, pfield.grid(), None, False, None, bbox, status, stats)
    final, epochs = _lib.grid_history(pfield.grid(), bbox)
    union = (bbox[:, 0].min(), bbox[:, 1].max(), bbox[:, 2].min(), bbox[:, 3].max())
    pfield.adopt(final, union)
    counts = np.zeros((pfield.nrows, pfield.ncols), dtype=np.int32)
    stats = _lib.TraceStats_()
    _launch(ctx, site, real, opts, pfield.grid(), epochs, True, counts, None, status, stats)
    return ChronicleResult(counts, status, stats, 2, bbox)


def warn_statuses(status):
    """Log one warning per unfinished trace, as compute_backtrace does (capturezone.py:249-250)."""
    bad = np.nonzero(status != _lib.TRACE_OK)[0]
    for k in bad[:100]:
        log.warning(' {0} (trace {1}, status {2}).'.format(
            STATUS_TEXT.get(int(status[k]), "trace ended abnormally"), int(k), int(status[k])))
    if bad.size > 100:
        log.warning(' ... {0} more unfinished traces.'.format(bad.size - 100))
