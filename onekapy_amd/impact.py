"""Impact curve, decile table and smoothed probability map of a
capture-probability field (oneka/visualize.py:228-233 and 354-397,
``create_probability_plot`` and ``create_impact_plot`` without the plots, and
the decile loop of oneka/oneka.py:279-288).

pr lists every node's Pr(capture) = pgrid / total_weight in descending order
and area[i] = (i + 1) * spacing^2 is the area where Pr(capture) >= pr[i].
With weight 1 per realization (stochastic.py:265) pgrid holds integer counts
k, so pr takes the values k / total_weight and the descending sort is a
counting sort: the GPU histogram of the counts (oka_count_histogram) gives the
same array as numpy's sort, value for value.  Any other field (weights other
than 1, e.g. compute_capturezone's weight argument) is sorted on the GPU by a
radix sort of the probabilities (oka_field_sorted_probabilities).
"""
import math

import numpy as np

from . import _lib

ACRE_M2 = 4046.86   # m^2 per acre (visualize.py:387, oneka.py:288)


class CaptureZoneError(Exception):
    """The field holds no realization (visualize.py)."""


def count_histogram(counts, max_count):
    """hist[k] = number of nodes whose count is k (0 <= k <= max_count), on the GPU."""
    counts = np.ascontiguousarray(counts, dtype=np.int32).ravel()
    hist = np.zeros(int(max_count) + 1, dtype=np.int64)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_count_histogram(ctx.h, counts.size, counts.ctypes.data, int(max_count),
                                          hist.ctypes.data, 0))
    return hist


def sorted_probabilities(pgrid, total_weight):
    """pgrid / total_weight in descending order, sorted on the GPU."""
    pgrid = np.ascontiguousarray(pgrid, dtype=np.float64).ravel()
    pr = np.empty_like(pgrid)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_field_sorted_probabilities(ctx.h, pgrid.size, pgrid.ctypes.data,
                                                     float(total_weight), pr.ctypes.data, 0))
    return pr


def impact_curve(spacing, pf):
    """(pr, area) as create_impact_plot returns them (visualize.py:382-385)."""
    if pf.total_weight <= 0:
        raise CaptureZoneError('Empty capture zone')
    counts = np.rint(pf.pgrid)
    if np.array_equal(counts, pf.pgrid) and counts.min(initial=0) >= 0 and \
            counts.max(initial=0) < 2 ** 31:
        top = int(counts.max(initial=0))
        hist = count_histogram(counts.astype(np.int32), top)
        values = np.arange(top, -1, -1, dtype=np.float64) / pf.total_weight   # descending
        pr = np.repeat(values, hist[::-1])
    else:
        pr = sorted_probabilities(pf.pgrid, pf.total_weight)
    n = pr.shape[0]
    area = (np.arange(n) + 1) * spacing ** 2
    return pr, area


def gaussian_weights(sigma, truncate=4.0):
    """The normalized symmetric kernel scipy.ndimage.gaussian_filter uses per
    axis for a zeroth-order filter: radius int(truncate sigma + 0.5) and
    weights exp(-x^2 / (2 sigma^2)) / their sum, computed with numpy in the
    same operations."""
    sd = float(sigma)
    radius = int(truncate * sd + 0.5)
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sd * sd) * x ** 2)
    return radius, phi / phi.sum()


def smoothed_probability(pf, smooth):
    """The probability map of create_probability_plot (visualize.py:228-233):
    Z = pgrid / total_weight, smoothed when smooth > 0 as
    scipy.ndimage.gaussian_filter(Z, smooth, mode='constant', cval=0.0) does
    (sigma in nodes), computed on the GPU.  Returns Z (nrows x ncols)."""
    if pf.total_weight <= 0:
        raise CaptureZoneError('Empty capture zone')
    pgrid = np.ascontiguousarray(pf.pgrid, dtype=np.float64)
    out = np.empty_like(pgrid)
    if not math.isfinite(float(smooth)):
        raise ValueError("smooth must be finite")
    radius, w = (0, None) if not smooth > 1e-15 else gaussian_weights(smooth)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_field_smooth(ctx.h, pgrid.shape[0], pgrid.shape[1], pgrid.ctypes.data,
                                       float(pf.total_weight), radius, _lib.ptr(w),
                                       out.ctypes.data, 0))
    return out


def decile_table(pr, area, levels=None):
    """Rows (level, pr[i], area [m^2], area [acres]) with i the first index
    whose pr <= level, for levels 0.05, 0.10, ..., 0.95 (oneka.py:284-288)."""
    if levels is None:
        levels = np.linspace(0.05, 0.95, 19)
    rows = []
    for p in levels:
        i = int(np.argmax(pr <= p))
        rows.append((float(p), float(pr[i]), float(area[i]), float(area[i] / ACRE_M2)))
    return rows
