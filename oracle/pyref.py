"""Reference-speed restatement of the capture-zone hot path in numpy-scalar Python.

TEST INFRASTRUCTURE ONLY.  It keeps the reference's cost structure: every
velocity evaluation is a Python loop over the wells on numpy float64 scalars,
every Dormand-Prince stage builds small numpy arrays, and rasterization is a
Python double loop over candidate nodes.  ``bench.py`` times it as the
reference-equivalent CPU baseline.  The tests use it to cross-check the C oracle
bit for bit: two independent restatements of the same reference lines.

Reference lines restated (OnekaPy tree):
  oneka/model.py:226-237, 258-266, 299-315, 341-350, 377-389, 423-427, 545-565, 598-604
  oneka/capturezone.py:109-123, 198-253
  oneka/probabilityfield.py:136-151, 196-261, 292-310, 334-339, 357-359, 407-427
  oneka/stochastic.py:220-265
"""
import math

import numpy as np

C4PI = 0.07957747154594767   # model.py:264
C2PI = 0.15915494309189535   # model.py:312-313

# capturezone.py:202-209: rows of the Dormand-Prince 5(4) tableau.
_DP_ROWS = (
    np.array([1 / 5]),
    np.array([3 / 40, 9 / 40]),
    np.array([44 / 45, -56 / 15, 32 / 9]),
    np.array([19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729]),
    np.array([9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656]),
)
_DP_B = np.array([35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84])
_DP_E = np.array([71 / 57600, -1 / 40, -71 / 16695, 71 / 1920, -17253 / 339200, 22 / 525])


class PyModel:
    """Per-realization analytic-element state (model.py:131-197)."""

    def __init__(self, base, conductivity, porosity, thickness, wells, xo=0.0, yo=0.0,
                 coef=None):
        self.base = base
        self.K = conductivity
        self.n = porosity
        self.b = thickness
        self.wells = [tuple(w) for w in wells]
        self.xo, self.yo = xo, yo
        self.coef = np.zeros(6) if coef is None else np.reshape(np.asarray(coef, float), 6)

    def wells_potential(self, x, y):
        acc = 0.0
        for xw, yw, _rw, qw in self.wells:
            ddx = x - xw
            ddy = y - yw
            acc += qw * np.log(ddx * ddx + ddy * ddy) * C4PI
        return acc

    def discharge(self, x, y):
        c = self.coef
        dx = x - self.xo
        dy = y - self.yo
        qx = -(2.0 * c[0] * dx + c[2] * dy + c[3])
        qy = -(2.0 * c[1] * dy + c[2] * dx + c[4])
        for xw, yw, _rw, qw in self.wells:
            ddx = x - xw
            ddy = y - yw
            r2 = ddx * ddx + ddy * ddy
            qx += -qw * ddx / r2 * C2PI
            qy += -qw * ddy / r2 * C2PI
        return qx, qy

    def velocity_confined(self, x, y):
        qx, qy = self.discharge(x, y)
        return qx / (self.b * self.n), qy / (self.b * self.n)

    def potential(self, x, y):
        """model.py:226-237: the regional quadratic plus the wells."""
        c = self.coef
        dx = x - self.xo
        dy = y - self.yo
        phi = c[0] * dx ** 2 + c[1] * dy ** 2 + c[2] * dx * dy + c[3] * dx + c[4] * dy + c[5]
        return phi + self.wells_potential(x, y)

    def velocity(self, x, y):
        """model.py:341-350 and 377-389: the unconfined velocity; raises
        ArithmeticError where the reference raises AquiferError."""
        qx, qy = self.discharge(x, y)
        phi = self.potential(x, y)
        if phi <= 0:
            raise ArithmeticError("potential <= 0")
        elif phi < 0.5 * self.K * self.b ** 2:
            head = np.sqrt(2.0 * phi / self.K)
        else:
            head = (phi + 0.5 * self.K * self.b ** 2) / (self.K * self.b)
        if head <= 0:
            raise ArithmeticError("head <= 0")
        elif head > self.b:
            return qx / (self.b * self.n), qy / (self.b * self.n)
        return qx / (head * self.n), qy / (head * self.n)

    def fit(self, obs, xo, yo):
        """model.py:497-606: weighted rows, then the reference's lstsq and inv."""
        nobs = len(obs)
        WA = np.zeros((nobs, 6))
        Wb = np.zeros((nobs, 1))
        for i, (x, y, z_ev, z_std) in enumerate(obs):
            h = z_ev - self.base
            if h >= self.b:
                pot_ev = self.K * self.b * (h - 0.5 * self.b)
                pot_std = self.K * self.b * z_std
            elif h > 0:
                pot_ev = 0.5 * self.K * (h ** 2 + z_std ** 2)
                pot_std = self.K * h * z_std
            else:
                raise ValueError("observation below the aquifer base")
            dx = x - xo
            dy = y - yo
            WA[i, :] = np.array([dx ** 2, dy ** 2, dx * dy, dx, dy, 1]) / pot_std
            Wb[i] = (pot_ev - self.wells_potential(x, y)) / pot_std
        ev = np.linalg.lstsq(WA, Wb, rcond=-1)[0]
        cov = np.linalg.inv(np.matmul(WA.T, WA))
        self.xo, self.yo = xo, yo
        self.coef = np.reshape(ev, 6)
        return ev, cov


def backtrace(xs, ys, duration, tol, maxstep, feval, max_steps=0):
    """capturezone.py:198-253.  Returns (vertices, attempted, accepted, status).

    Deliberate deviations, both also in the C oracle and the GPU path: a NaN or
    zero step (where the reference would loop forever) ends the trace with
    status 2, and an optional attempted-step cap ends it with status 3.
    """
    eps = np.finfo(float).eps
    t = 0
    dt = 0.1 * np.sign(duration)
    xy = np.array([xs, ys])
    verts = [(xs, ys)]
    attempted = accepted = 0
    status = 0
    r2, r3, r4, r5, r6 = _DP_ROWS
    b, e = _DP_B, _DP_E
    try:
        k1 = feval(xy)
        while abs(t) < abs(duration):
            if max_steps and attempted >= max_steps:
                status = 3
                break
            if abs(t + dt) > abs(duration):
                dt = duration - t
            attempted += 1
            k2 = feval(xy + dt * (r2[0] * k1))
            k3 = feval(xy + dt * (r3[0] * k1 + r3[1] * k2))
            k4 = feval(xy + dt * (r4[0] * k1 + r4[1] * k2 + r4[2] * k3))
            k5 = feval(xy + dt * (r5[0] * k1 + r5[1] * k2 + r5[2] * k3 + r5[3] * k4))
            k6 = feval(xy + dt * (r6[0] * k1 + r6[1] * k2 + r6[2] * k3 + r6[3] * k4
                                  + r6[4] * k5))
            xyn = xy + dt * (b[0] * k1 + b[2] * k3 + b[3] * k4 + b[4] * k5 + b[5] * k6)
            k7 = feval(xyn)
            est = np.linalg.norm(dt * (e[0] * k1 + e[1] * k7 + e[2] * k3 + e[3] * k4
                                       + e[4] * k5 + e[5] * k6), np.inf)
            ds = np.linalg.norm(xyn - xy)
            if est < tol and ds < maxstep:
                t = t + dt
                k1 = k7
                xy = xyn
                verts.append(xy)
                accepted += 1
            dt = 0.9 * min((tol / (est + eps)) ** (1 / 5), maxstep / (ds + eps), 10) * dt
            if abs(t) < abs(duration) and (not (dt != 0) or dt != dt):
                status = 2
                break
    except ArithmeticError:
        status = 1
    return verts, attempted, accepted, status


class PyField:
    """probabilityfield.py restated: node lattice, per-realization marks."""

    def __init__(self, dx, dy, xo, yo):
        self.dx, self.dy = dx, dy
        self.xmin, self.xmax = xo - dx, xo + dx
        self.ymin, self.ymax = yo - dy, yo + dy
        self.nrows = self.ncols = 3
        self.pgrid = np.zeros((3, 3))
        self.rgrid = np.zeros((3, 3), dtype=bool)
        self.total_weight = 0.0

    def expand(self, xmin, xmax, ymin, ymax):
        nr, nc = self.nrows, self.ncols
        rs = cs = 0
        while xmin <= self.xmin:
            self.xmin -= self.dx
            nc += 1
            cs += 1
        while xmax >= self.xmax:
            self.xmax += self.dx
            nc += 1
        while ymin <= self.ymin:
            self.ymin -= self.dy
            nr += 1
            rs += 1
        while ymax >= self.ymax:
            self.ymax += self.dy
            nr += 1
        if (nr, nc) != (self.nrows, self.ncols):
            pg = np.zeros((nr, nc))
            rg = np.zeros((nr, nc), dtype=bool)
            pg[rs:rs + self.nrows, cs:cs + self.ncols] = self.pgrid
            rg[rs:rs + self.nrows, cs:cs + self.ncols] = self.rgrid
            self.pgrid, self.rgrid = pg, rg
            self.nrows, self.ncols = nr, nc

    @staticmethod
    def dist2(ax, ay, bx, by, cx, cy):
        bax, bay = bx - ax, by - ay
        cax, cay = cx - ax, cy - ay
        perp = bax * cay - bay * cax
        dot = bax * cax + bay * cay
        l2 = bax * bax + bay * bay
        a2 = perp * perp / l2
        b2 = dot * dot / l2
        if dot < 0:
            return a2 + b2
        if b2 > l2:
            return a2 + b2 - 2 * dot + l2
        return a2

    def insert(self, ax, ay, bx, by, umbra):
        u2 = umbra * umbra
        j0 = max(0, math.floor((min(ax, bx) - umbra - self.xmin) / self.dx))
        j1 = min(self.ncols, math.floor((max(ax, bx) + umbra - self.xmin) / self.dx) + 1)
        i0 = max(0, math.floor((min(ay, by) - umbra - self.ymin) / self.dy))
        i1 = min(self.nrows, math.floor((max(ay, by) + umbra - self.ymin) / self.dy) + 1)
        for j in np.arange(j0, j1):
            for i in np.arange(i0, i1):
                if not self.rgrid[i, j]:
                    if self.dist2(ax, ay, bx, by, self.xmin + j * self.dx,
                                  self.ymin + i * self.dy) < u2:
                        self.rgrid[i, j] = True

    def rasterize(self, x, y, umbra):
        self.expand(min(x), max(x), min(y), max(y))
        for k in range(len(x) - 1):
            self.insert(x[k], y[k], x[k + 1], y[k + 1], umbra)

    def register(self, w):
        self.total_weight += w
        self.pgrid[self.rgrid] += w
        self.rgrid[:] = False


def start_points(xt, yt, rt, npaths):
    """capturezone.py:110-115 (STEPAWAY = 1 m)."""
    th = np.linspace(0, 2 * np.pi, npaths + 1)[0:-1]
    return [((rt + 1.0) * np.cos(a) + xt, (rt + 1.0) * np.sin(a) + yt) for a in th]


def realization(model, starts, duration, tol, maxstep, field, umbra, weight=1.0):
    """capturezone.py:113-123 with the confined feval of stochastic.py:253-256."""
    def feval(xy):
        vx, vy = model.velocity_confined(xy[0], xy[1])
        return np.array([-vx, -vy])

    steps = [0, 0]
    for xs, ys in starts:
        verts, att, acc, _st = backtrace(xs, ys, duration, tol, maxstep, feval)
        steps[0] += att
        steps[1] += acc
        field.rasterize([v[0] for v in verts], [v[1] for v in verts], umbra)
    field.register(weight)
    return steps
