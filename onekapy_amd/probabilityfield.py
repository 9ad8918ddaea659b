"""Auto-expanding, axis-aligned node lattice (oneka/probabilityfield.py:53-427).

``ProbabilityField`` keeps the reference's public attributes (deltax, deltay,
xmin/xmax/ymin/ymax, nrows/ncols, pgrid float64, rgrid bool, total_weight) and
methods.  The grid bookkeeping (``expand``, ``register``, ``reset``) is host
code with the reference's exact update rules; the vector-to-raster work
(``insert``, ``rasterize``) runs in liboka's FP64 raster kernel.
"""
import logging
import math

import numpy as np

from . import _lib

log = logging.getLogger("Oneka")


class Error(Exception):
    """Base class for module errors (probabilityfield.py:41-43)."""


class RangeError(Error):
    """Passed argument out of valid range (probabilityfield.py:47-49)."""


class ProbabilityField:
    """An auto-expanding, axis-aligned-grid-based probability field.

    Node (i, j) is at (xmin + j*deltax, ymin + i*deltay); row i is y.
    """

    def __init__(self, deltax, deltay, xo=np.nan, yo=np.nan):
        # probabilityfield.py:125-151
        if deltax <= 0:
            raise RangeError("<deltax> must be > 0.")
        if deltay <= 0:
            raise RangeError("<deltay> must be > 0.")
        self.deltax = deltax
        self.deltay = deltay
        if np.isnan(xo) or np.isnan(yo):
            self.nrows = 0
            self.ncols = 0
        else:
            self.xmin = xo - deltax
            self.xmax = xo + deltax
            self.ymin = yo - deltay
            self.ymax = yo + deltay
            self.nrows = 3
            self.ncols = 3
            self.pgrid = np.zeros((self.nrows, self.ncols), dtype=np.float64)
            self.rgrid = np.zeros((self.nrows, self.ncols), dtype=bool)
            self.total_weight = 0.0

    def __repr__(self):
        if self.nrows == 0:
            return "ProbabilityField({0.deltax}, {0.deltay})".format(self)
        return ("ProbabilityField({0.deltax}, {0.deltay}, {0.nrows}, {0.ncols}, "
                "{0.xmin}, {0.xmax}, {0.ymin}, {0.ymax}, {0.total_weight})".format(self))

    def __str__(self):
        if self.nrows == 0:
            return "ProbabilityField({0.deltax}, {0.deltay}, {0.nrows}, {0.ncols})".format(self)
        return ("ProbabilityField({0.deltax}, {0.deltay}, {0.nrows}, {0.ncols}, "
                "{0.xmin}, {0.xmax}, {0.ymin}, {0.ymax}, {0.total_weight}, "
                "\n{0.pgrid!r}, \n{0.rgrid!r})".format(self))

    @property
    def empty(self):
        return self.ncols == 0 or self.nrows == 0

    def expand(self, xmin, xmax, ymin, ymax):
        """Grow the grids so that the bbox lies strictly inside (probabilityfield.py:175-261)."""
        if xmin > xmax:
            raise RangeError("<xmin> must be <= <xmax>.")
        if ymin > ymax:
            raise RangeError("<ymin> must be <= <ymax>.")
        if self.empty:
            self.xmin = xmin - self.deltax
            self.ymin = ymin - self.deltay
            self.ncols = max(3, math.ceil((xmax - self.xmin) / self.deltax) + 2)
            self.nrows = max(3, math.ceil((ymax - self.ymin) / self.deltay) + 2)
            self.xmax = self.xmin + (self.ncols - 1) * self.deltax
            self.ymax = self.ymin + (self.nrows - 1) * self.deltay
            self.pgrid = np.zeros((self.nrows, self.ncols), dtype=np.float64)
            self.rgrid = np.zeros((self.nrows, self.ncols), dtype=bool)
            self.total_weight = 0.0
            log.debug('grids expanded to ({0}, {1}).'.format(self.nrows, self.ncols))
            return
        nrows, ncols = self.nrows, self.ncols
        rshift = cshift = 0
        while xmin <= self.xmin:
            self.xmin -= self.deltax
            ncols += 1
            cshift += 1
        while xmax >= self.xmax:
            self.xmax += self.deltax
            ncols += 1
        while ymin <= self.ymin:
            self.ymin -= self.deltay
            nrows += 1
            rshift += 1
        while ymax >= self.ymax:
            self.ymax += self.deltay
            nrows += 1
        if (nrows != self.nrows) or (ncols != self.ncols):
            pgrid = np.zeros((nrows, ncols), dtype=np.float64)
            rgrid = np.zeros((nrows, ncols), dtype=bool)
            rgrid[rshift:rshift + self.nrows, cshift:cshift + self.ncols] = self.rgrid
            pgrid[rshift:rshift + self.nrows, cshift:cshift + self.ncols] = self.pgrid
            self.rgrid = rgrid
            self.pgrid = pgrid
            self.nrows = nrows
            self.ncols = ncols
            log.debug(' grids expanded to ({0}, {1})'.format(nrows, ncols))

    # -- GPU rasterization ----------------------------------------------------
    def grid(self):
        """This field's lattice as an oka_grid (an all-zero grid when empty)."""
        if self.empty:
            g = _lib.Grid_()
            g.dx, g.dy = float(self.deltax), float(self.deltay)
            return g
        return _lib.make_grid(self.xmin, self.ymin, self.xmax, self.ymax, self.deltax,
                              self.deltay, self.nrows, self.ncols)

    def _mark(self, xy, umbra):
        xy = _lib.f64(xy).reshape(-1, 2)
        if xy.shape[0] < 2 or self.empty:
            return
        marks = np.ascontiguousarray(self.rgrid, dtype=np.uint8)
        ctx = _lib.context()
        g = self.grid()
        ctx.check(ctx.lib.oka_rasterize(ctx.h, g, xy.shape[0], _lib.ptr(xy), float(umbra),
                                        _lib.ptr(marks), 0))
        self.rgrid = marks.astype(bool)

    def insert(self, ax, ay, bx, by, umbra):
        """Mark nodes strictly within umbra of one segment (probabilityfield.py:264-310)."""
        self._mark([[ax, ay], [bx, by]], umbra)

    def rasterize(self, x, y, umbra):
        """Expand to the track's bbox, then mark its segments (probabilityfield.py:313-339)."""
        self.expand(min(x), max(x), min(y), max(y))
        self._mark(np.stack([np.asarray(x, dtype=np.float64),
                             np.asarray(y, dtype=np.float64)], axis=1), umbra)

    def register(self, weight):
        """Add this realization's marks with weight, then clear them (probabilityfield.py:342-359)."""
        self.total_weight += weight
        self.pgrid[self.rgrid] += weight
        self.rgrid[:] = False

    def reset(self):
        """Clear the registration grid (probabilityfield.py:362-376)."""
        self.rgrid[:] = False

    @staticmethod
    def distancesquared(ax, ay, bx, by, cx, cy):
        """Squared distance from (cx, cy) to the segment [a, b] (probabilityfield.py:379-427)."""
        bax = bx - ax
        bay = by - ay
        cax = cx - ax
        cay = cy - ay
        perpdot = bax * cay - bay * cax
        dot = bax * cax + bay * cay
        length2 = bax * bax + bay * bay
        alpha2 = perpdot * perpdot / length2
        beta2 = dot * dot / length2
        if dot < 0:
            d2 = alpha2 + beta2
        elif beta2 > length2:
            d2 = alpha2 + beta2 - 2 * dot + length2
        else:
            d2 = alpha2
        return d2

    # -- used by the batched paths ---------------------------------------------
    def adopt(self, final, union):
        """Take the geometry of `final`, the oka_grid the library computed for the
        expansion history of a batch of traces whose overall bbox is `union`
        (xmin, xmax, ymin, ymax), keeping the accumulated content."""
        if self.empty:
            self.xmin, self.ymin = final.xmin, final.ymin
            self.xmax, self.ymax = final.xmax, final.ymax
            self.nrows, self.ncols = int(final.nrows), int(final.ncols)
            self.pgrid = np.zeros((self.nrows, self.ncols), dtype=np.float64)
            self.rgrid = np.zeros((self.nrows, self.ncols), dtype=bool)
            self.total_weight = 0.0
            return
        # Each growth loop of expand() only moves outwards, so the state after
        # a sequence of expands equals the state after one expand with the union
        # of their boxes.
        self.expand(*union)
        if (self.nrows, self.ncols) != (final.nrows, final.ncols) or \
                self.xmin != final.xmin or self.ymin != final.ymin or \
                self.xmax != final.xmax or self.ymax != final.ymax:
            raise RuntimeError("expansion history and expand() disagree")
