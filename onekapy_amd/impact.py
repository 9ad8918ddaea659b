"""Impact curve and decile table of a capture-probability field
(oneka/visualize.py:345-390 ``create_impact_plot`` without the plot, and the
decile loop of oneka/oneka.py:279-288).

pr lists every node's Pr(capture) = pgrid / total_weight in descending order
and area[i] = (i + 1) * spacing^2 is the area where Pr(capture) >= pr[i].
With weight 1 per realization (stochastic.py:265) pgrid holds integer counts
k, so pr takes the values k / total_weight and the descending sort is a
counting sort: the GPU histogram of the counts (oka_count_histogram) gives the
same array as numpy's sort, value for value.
"""
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


def impact_curve(spacing, pf):
    """(pr, area) as create_impact_plot returns them (visualize.py:382-385).

    The field's pgrid must hold whole counts (weight 1 per realization, as
    every capture zone of this package and of the reference has)."""
    if pf.total_weight <= 0:
        raise CaptureZoneError('Empty capture zone')
    counts = np.rint(pf.pgrid)
    if not np.array_equal(counts, pf.pgrid) or counts.min(initial=0) < 0:
        raise ValueError("pgrid does not hold whole counts (weights other than 1)")
    top = int(counts.max(initial=0))
    hist = count_histogram(counts.astype(np.int32), top)
    values = np.arange(top, -1, -1, dtype=np.float64) / pf.total_weight   # descending
    pr = np.repeat(values, hist[::-1])
    n = pr.shape[0]
    area = (np.arange(n) + 1) * spacing ** 2
    return pr, area


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
