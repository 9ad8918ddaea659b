"""Learned control variate for the capture probability.

No reference counterpart: the reference estimates each node's capture
probability as the fraction of traced realizations whose traces come within
umbra of it (stochastic.py:220-265, probabilityfield.py:342-359).  Here a
surrogate network s(theta) predicts, from a realization's parameters theta
(discharges, conductivity, porosity, thickness, A..F), the fraction of each
window cell's nodes that the realization visits, and serves as a control
variate:

    p_hat(c) = mean over M surrogate-only draws of s_c
             + mean over N traced realizations of (y_c - s_c)

which is unbiased for the cell capture probability p(c) = E[y_c] whatever the
surrogate, and whose variance Var(y - s)/N + Var(s)/M is far below the plain
Var(y)/N when s is a good predictor.  The N traced realizations are used both
to train and to correct by K-fold cross-fitting (each fold is corrected by a
network that never saw it), so no traced realization is spent on training
alone.  The plain tracer remains the fallback and the reference for the error.

At node resolution (control_variate_nodes, the ProbabilityField the
reference returns) one network trained on a separate set of traced
realizations serves every node of the window: node j's estimate is its plain
mean over the correction realizations minus a linear combination of the
surrogate outputs of the 3 x 3 cells around its own (each less its mean over
fresh surrogate draws), with coefficients fitted on the other part of the
correction set, so that the estimate stays unbiased.  Nodes outside the
window keep the plain estimate of all traced realizations.

The network (n_in -> 512 -> 512 -> 4096, GELU, BF16 weights, FP32
accumulation) is trained and evaluated by liboka's hand-written BF16 MFMA
kernels (oka_mlp_*).
"""
import ctypes as C
import math
import threading
import time

import numpy as np

from . import _lib
from .capturezone import start_points
from .stochastic import draw_realizations, fit_and_draw


class SurrogateMLP:
    """n_in -> n_hidden -> n_hidden -> n_out perceptron on the GPU (oka_mlp_*)."""

    def __init__(self, n_in, n_hidden=512, n_out=4096, max_batch=8192, seed=0, device=0,
                 ctx=None):
        """ctx: the liboka context (and so the stream) the network lives on;
        default the process-wide context of `device`."""
        self.ctx = ctx if ctx is not None else _lib.context(device)
        self.n_in, self.n_hidden, self.n_out = int(n_in), int(n_hidden), int(n_out)
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.oka_mlp_create(self.ctx.h, self.n_in, self.n_hidden,
                                                   self.n_out, self.max_batch, C.byref(h)))
        self.h = h
        self.n_params = int(self.ctx.lib.oka_mlp_param_count(h))
        self.set_params(self.initial_params(seed))

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.oka_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def shapes(self):
        I, H, O = self.n_in, self.n_hidden, self.n_out
        return [("W1", (H, I)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (O, H)),
                ("b3", (O,))]

    def split(self, flat):
        """The flat parameter vector as a dict of arrays (views)."""
        out, k = {}, 0
        for name, shp in self.shapes():
            n = int(np.prod(shp))
            out[name] = flat[k:k + n].reshape(shp)
            k += n
        return out

    def initial_params(self, seed):
        """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases."""
        rng = np.random.default_rng(seed)
        parts = []
        fan = {"W1": self.n_in, "b1": self.n_in, "W2": self.n_hidden, "b2": self.n_hidden,
               "W3": self.n_hidden, "b3": self.n_hidden}
        for name, shp in self.shapes():
            bound = 1.0 / math.sqrt(fan[name])
            parts.append(rng.uniform(-bound, bound, shp).astype(np.float32).ravel())
        return np.concatenate(parts)

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.n_params
        self.ctx.check(self.ctx.lib.oka_mlp_set_params(self.h, flat.ctypes.data, 0))

    def get_params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.oka_mlp_get_params(self.h, out.ctypes.data, 0))
        return out

    def grad(self, x, marks, target_scale):
        """(gradient, loss) of the mean BCE on one batch."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        marks = np.ascontiguousarray(marks, dtype=np.uint8)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float(0.0)
        self.ctx.check(self.ctx.lib.oka_mlp_grad(self.h, x.shape[0], x.ctypes.data,
                                                 marks.ctypes.data, float(target_scale),
                                                 g.ctypes.data, C.byref(loss), 0))
        return g, loss.value

    def train(self, x, marks, target_scale, epochs, batch=512, lr=1e-3, beta1=0.9, beta2=0.999,
              eps=1e-8, seed=0, rows=None):
        """Adam over shuffled minibatches of the samples `rows` (default: all);
        returns the per-step losses."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        marks = np.ascontiguousarray(marks, dtype=np.uint8)
        rows = np.arange(x.shape[0]) if rows is None else np.asarray(rows)
        n = rows.shape[0]
        per_epoch = n // batch
        if per_epoch == 0:
            raise ValueError("fewer samples than one batch")
        rng = np.random.default_rng(seed)
        # every epoch's shuffle up front: the whole schedule is one library call
        order = np.concatenate([rows[rng.permutation(n)[:per_epoch * batch]]
                                for _ in range(int(epochs))]).astype(np.int32)
        n = x.shape[0]
        nsteps = int(epochs) * per_epoch
        losses = np.zeros(nsteps, dtype=np.float32)
        self.ctx.check(self.ctx.lib.oka_mlp_train(
            self.h, n, x.ctypes.data, marks.ctypes.data, float(target_scale), order.ctypes.data,
            batch, nsteps, lr, beta1, beta2, eps, losses.ctypes.data, 0))
        return losses

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        probs = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        self.ctx.check(self.ctx.lib.oka_mlp_predict(self.h, x.shape[0], x.ctypes.data,
                                                    probs.ctypes.data, 0))
        return probs

    def accumulate(self, x, marks=None, target_scale=1.0):
        """Sums over samples of s, s^2 and (with marks) y - s, (y - s)^2 per output."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        acc = {k: np.zeros(self.n_out) for k in ("s", "s2", "d", "d2")}
        mp = None
        if marks is not None:
            marks = np.ascontiguousarray(marks, dtype=np.uint8)
            mp = marks.ctypes.data
        self.ctx.check(self.ctx.lib.oka_mlp_accumulate(
            self.h, x.shape[0], x.ctypes.data, mp, float(target_scale), acc["s"].ctypes.data,
            acc["s2"].ctypes.data, acc["d"].ctypes.data if mp else None,
            acc["d2"].ctypes.data if mp else None, 0))
        return acc

    def surrogate_sums(self, site, n, seed, first, std, antithetic=False, qmc=False):
        """Sums of s and s^2 over n fresh realizations drawn, fitted and
        featurized on the device (oka_cv_surrogate_sums).  With antithetic
        the realizations come in reflected pairs and "s2" sums the squared
        pair means (n / 2 of them); with qmc they are Sobol' points
        first .. first + n - 1 under the nested scramble keyed by `seed`."""
        from .stochastic import oka_dists
        acc = {"s": np.zeros(self.n_out), "s2": np.zeros(self.n_out)}
        wxy = site.wells_xy()
        obs = site.obs_array()
        s = _lib.make_site(wxy, site.xt, site.yt, site.base, obs)
        dists = oka_dists(site.wells, site.c_dist, site.p_dist, site.t_dist)
        mean = np.ascontiguousarray(std.mean, dtype=np.float64)
        sd = np.ascontiguousarray(std.std, dtype=np.float64)
        self.ctx.check(self.ctx.lib.oka_cv_surrogate_sums(
            self.ctx.h, self.h, s, C.addressof(dists), int(n), int(seed) & (2 ** 64 - 1),
            int(first), mean.ctypes.data, sd.ctypes.data, acc["s"].ctypes.data,
            acc["s2"].ctypes.data,
            (_lib.ANTITHETIC if antithetic else 0) | (_lib.QMC if qmc else 0)))
        if not np.all(np.isfinite(acc["s"])):
            raise ArithmeticError("the fit failed for some surrogate draws")
        return acc


def _surrogate_sums_qmc(net, site, npts, seeds, std):
    """Row k: the surrogate sums of Sobol' points 0 .. npts - 1 under the
    scramble of seeds[k] (oka_cv_surrogate_sums_qmc, one synchronization)."""
    from .stochastic import oka_dists
    seeds = np.ascontiguousarray([int(x) & (2 ** 64 - 1) for x in seeds], dtype=np.uint64)
    out = np.zeros((len(seeds), net.n_out))
    wxy = site.wells_xy()
    obs = site.obs_array()
    s = _lib.make_site(wxy, site.xt, site.yt, site.base, obs)
    dists = oka_dists(site.wells, site.c_dist, site.p_dist, site.t_dist)
    mean = np.ascontiguousarray(std.mean, dtype=np.float64)
    sd = np.ascontiguousarray(std.std, dtype=np.float64)
    net.ctx.check(net.ctx.lib.oka_cv_surrogate_sums_qmc(
        net.ctx.h, net.h, s, C.addressof(dists), int(npts), len(seeds), seeds.ctypes.data,
        mean.ctypes.data, sd.ctypes.data, out.ctypes.data, 0))
    if not np.all(np.isfinite(out)):
        raise ArithmeticError("the fit failed for some surrogate draws")
    return out


def features(d, coef):
    """Per-realization parameters: discharges, K, n, b, A..F (nwells + 9 columns)."""
    return np.concatenate([d["q"], d["K"][:, None], d["n"][:, None], d["b"][:, None],
                           np.asarray(coef)], axis=1)


class Standardizer:
    def __init__(self, x):
        self.mean = x.mean(0)
        self.std = x.std(0)
        self.std[self.std == 0] = 1.0

    def __call__(self, x):
        return ((x - self.mean) / self.std).astype(np.float32)


def _site_struct(site):
    wxy = site.wells_xy()
    return wxy, _lib.make_site(wxy, site.xt, site.yt)


def trace_counts(site, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_AUTO, far_order=0):
    """Fine-grid counts of a batch of realizations on the site's fixed grid."""
    pf = site.fixed_field()
    g = pf.grid()
    wxy, s = _site_struct(site)
    starts = start_points(site.xt, site.yt, site.rt, site.npaths)
    o = _lib.make_trace_opts(site.npaths, site.duration, site.tol, site.maxstep, site.umbra,
                             _lib.ptr(starts), precision, site.confined, schedule=schedule,
                             far_order=far_order)
    counts = np.zeros((pf.nrows, pf.ncols), np.int32)
    st = _lib.TraceStats_()
    K, n, b, q = (np.ascontiguousarray(d[k]) for k in ("K", "n", "b", "q"))
    coef = np.ascontiguousarray(coef)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_trace_chronicle(ctx.h, s, len(K), _lib.ptr(K), _lib.ptr(n), _lib.ptr(b),
                                          _lib.ptr(q), _lib.ptr(coef), o, g, 0, None, 1,
                                          _lib.ptr(counts), None, None, None, st, 0))
    return counts, st


def choose_window(counts, cells=64, margin=1.1):
    """A cells x cells window of square cells covering the visited nodes."""
    nz = np.argwhere(counts > 0)
    if nz.size == 0:
        r = counts.shape[0] // 2
        c = counts.shape[1] // 2
        return _window(r - cells // 2, c - cells // 2, 1, cells)
    r0, r1 = nz[:, 0].min(), nz[:, 0].max()
    c0, c1 = nz[:, 1].min(), nz[:, 1].max()
    ext = max(r1 - r0 + 1, c1 - c0 + 1)
    cell = max(1, min(15, int(math.ceil(ext * margin / cells))))
    rc, cc = (r0 + r1) // 2, (c0 + c1) // 2
    return _window(int(rc - cells // 2 * cell), int(cc - cells // 2 * cell), cell, cells)


def _window(row0, col0, cell, cells):
    w = _lib.Window_()
    w.row0, w.col0, w.cell, w.nrows, w.ncols = row0, col0, cell, cells, cells
    return w


def trace_marks(site, d, coef, window, precision=_lib.FP32, masks=False, counts=None):
    """Per-realization visited-node counts per window cell (uint8 [R, cells]);
    with masks, also which of the cell's nodes (uint16 [R, cells], bit
    i * cell + j for fine node (i, j) of the cell; cells of at most 4 x 4
    nodes).  counts (int32 [nrows, ncols] of the site's fixed grid, or None)
    accumulates every realization's visited nodes."""
    pf = site.fixed_field()
    g = pf.grid()
    wxy, s = _site_struct(site)
    starts = start_points(site.xt, site.yt, site.rt, site.npaths)
    o = _lib.make_trace_opts(site.npaths, site.duration, site.tol, site.maxstep, site.umbra,
                             _lib.ptr(starts), precision, site.confined)
    R = len(d["K"])
    marks = np.zeros((R, window.nrows * window.ncols), np.uint8)
    mk = np.zeros((R, window.nrows * window.ncols), np.uint16) if masks else None
    if counts is not None and (counts.dtype != np.int32 or not counts.flags.c_contiguous or
                               counts.shape != (pf.nrows, pf.ncols)):
        raise ValueError("counts must be a C-contiguous int32 array of the grid's shape")
    st = _lib.TraceStats_()
    K, n, b, q = (np.ascontiguousarray(d[k]) for k in ("K", "n", "b", "q"))
    coef = np.ascontiguousarray(coef)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_trace_marks(ctx.h, s, R, _lib.ptr(K), _lib.ptr(n), _lib.ptr(b),
                                      _lib.ptr(q), _lib.ptr(coef), o, g, window, _lib.ptr(marks),
                                      _lib.ptr(mk), _lib.ptr(counts), st, 0))
    return (marks, mk, st) if masks else (marks, st)


def cell_probabilities(counts, nreal, window):
    """Fine-grid counts -> the window cells' capture probability (mean over each
    cell's nodes of count / nreal; nodes outside the grid count as 0)."""
    c, nr, nc = window.cell, window.nrows, window.ncols
    out = np.zeros((nr * c, nc * c))
    r0, c0 = window.row0, window.col0
    gr0, gc0 = max(r0, 0), max(c0, 0)
    gr1 = min(r0 + nr * c, counts.shape[0])
    gc1 = min(c0 + nc * c, counts.shape[1])
    if gr1 > gr0 and gc1 > gc0:
        out[gr0 - r0:gr1 - r0, gc0 - c0:gc1 - c0] = counts[gr0:gr1, gc0:gc1]
    return out.reshape(nr, c, nc, c).mean(axis=(1, 3)).ravel() / float(nreal)


def plain_estimate(marks, tscale):
    """Mean of marks * tscale per cell and the variance of that mean, from the
    exact integer moments of the marks (oka_mark_moments)."""
    marks = np.ascontiguousarray(marks, dtype=np.uint8)
    n, ncell = marks.shape
    if n == 0:
        raise ValueError("no realizations")
    s1 = np.zeros(ncell, np.uint64)
    s2 = np.zeros(ncell, np.uint64)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_mark_moments(ctx.h, n, ncell, marks.ctypes.data, s1.ctypes.data,
                                       s2.ctypes.data, 0))
    mean = s1.astype(np.float64) / n
    var = np.maximum(s2.astype(np.float64) / n - mean ** 2, 0.0)
    return mean * tscale, var * (tscale * tscale) / n


class ControlVariateResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def qmc_split(m, replicates, pow2=True):
    """(points per replicate, replicates) for m QMC evaluations: m // replicates
    points, or with pow2 the largest power of two not above that (a Sobol' net's
    balanced size) and as many replicates as then fit in m."""
    npts = m // replicates
    if npts < 1:
        raise ValueError(f"{m} QMC evaluations cannot fill {replicates} replicates")
    if pow2:
        npts = 1 << (npts.bit_length() - 1)
        return npts, m // npts
    return npts, replicates


def _replicate_seed(base, fold, k):
    """Seed of fold `fold`'s k-th QMC replicate (an independent scramble)."""
    return int(np.random.SeedSequence([base, fold, k]).generate_state(1, np.uint64)[0])


def control_variate(site, n_traced, n_surrogate, folds=2, train_steps=2000, batch=512, lr=1e-3,
                    seed=0, precision=_lib.FP32, pilot=2000, window=None, n_hidden=512,
                    antithetic=False, qmc_replicates=0, qmc_pow2=False):
    """Cross-fitted control-variate estimate of the window cells' capture probability.

    Each fold's network takes train_steps Adam steps of `batch` samples (whole
    shuffled epochs, rounded up).  With antithetic, the surrogate term draws
    its realizations in antithetic pairs and its variance is that of the pair
    means over the number of pairs.  With qmc_replicates = R > 0 instead, each
    fold's surrogate realizations are R independently scrambled copies of the
    first M_f / R Sobol' points (randomized quasi-Monte Carlo; with qmc_pow2
    the points per copy are rounded down to a power of two and R raised to
    fill M_f, see qmc_split); the surrogate mean is the mean of the replicate
    means and its variance their sample variance over the replicate count.

    Returns a ControlVariateResult with p_cv, its estimated variance var_cv,
    the plain estimate p_plain from the same traced realizations with its
    variance var_plain, the window, and the time of each phase (seconds).
    The folds run concurrently: times["folds"] is their wall time, "train"
    and "surrogate" the longest fold's time in each of its two phases.
    """
    if antithetic and qmc_replicates:
        raise ValueError("antithetic pairs and QMC replicates exclude each other")
    t = {}
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    d = draw_realizations(site.wells, site.c_dist, site.p_dist, site.t_dist, n_traced, rng)
    wxy = site.wells_xy()
    _, _, coef = fit_and_draw(wxy, site.xt, site.yt, site.base, site.observations, d["K"],
                              d["b"], d["q"], d["z"], moments=False)
    t["fit_draw"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    if window is None:
        npil = min(pilot, n_traced)
        pc, _ = trace_counts(site, {k: v[:npil] for k, v in d.items()}, coef[:npil], precision)
        window = choose_window(pc)
    t["pilot"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    marks, st = trace_marks(site, d, coef, window, precision)
    t["trace"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    tscale = 1.0 / (window.cell * window.cell)
    F = features(d, coef)
    std = Standardizer(F)
    X = std(F)
    # folds are contiguous slices: the realizations are independent draws
    bounds = np.linspace(0, n_traced, folds + 1).astype(np.int64)
    surr_seed = int(rng.integers(1, 2 ** 63))
    surr_split = np.linspace(0, n_surrogate, folds + 1).astype(np.int64)
    if antithetic:   # pairs (2k, 2k + 1) stay within one fold
        surr_split = surr_split // 2 * 2
    ncell = window.nrows * window.ncols
    # one network per fold, each on its own context and stream: the folds'
    # training and surrogate passes are independent and run concurrently
    # (one host thread per context), which fills the GPU better than the
    # small launches of one network alone
    ctxs = [_lib.context()] + [_lib.Context(_lib.context().device) for _ in range(folds - 1)]
    nets = [SurrogateMLP(X.shape[1], n_hidden, ncell, seed=seed + 1 + f, ctx=ctxs[f])
            for f in range(folds)]
    t["setup"] = time.perf_counter() - t0
    out = [None] * folds

    def run_fold(f):
        lo, hi = int(bounds[f]), int(bounds[f + 1])
        tr = np.concatenate([np.arange(0, lo), np.arange(hi, n_traced)])
        t1 = time.perf_counter()
        epochs = max(1, int(math.ceil(train_steps / max(1, len(tr) // batch))))
        loss = nets[f].train(X, marks, tscale, epochs, batch, lr, seed=seed + 100 + f, rows=tr)
        t2 = time.perf_counter()
        mf = int(surr_split[f + 1] - surr_split[f])
        if qmc_replicates:
            npts, nrep = qmc_split(mf, qmc_replicates, qmc_pow2)
            mf = npts * nrep
            seeds = [_replicate_seed(surr_seed, f, k) for k in range(nrep)]
            a_s = {"reps": _surrogate_sums_qmc(nets[f], site, npts, seeds, std) / npts}
        else:
            a_s = nets[f].surrogate_sums(site, mf, surr_seed, int(surr_split[f]), std,
                                         antithetic)
        a_d = nets[f].accumulate(X[lo:hi], marks[lo:hi], tscale)
        out[f] = (loss, a_s, a_d, mf, hi - lo, t2 - t1, time.perf_counter() - t2)

    t0 = time.perf_counter()
    errors = []

    def guarded(f):
        try:
            run_fold(f)
        except BaseException as e:   # re-raised in the calling thread
            errors.append(e)

    threads = [threading.Thread(target=guarded, args=(f,)) for f in range(1, folds)]
    for th in threads:
        th.start()
    guarded(0)
    for th in threads:
        th.join()
    if errors:
        raise errors[0]
    t["folds"] = time.perf_counter() - t0
    t["train"] = max(o[5] for o in out)
    t["surrogate"] = max(o[6] for o in out)
    sum_s = np.zeros(ncell)
    var_cv = np.zeros(ncell)
    sum_d = np.zeros(ncell)
    losses = []
    for f in range(folds):
        loss, a_s, a_d, mf, nf, _, _ = out[f]
        losses.append(loss)
        w = nf / n_traced
        if qmc_replicates:
            # replicate means are independent and unbiased: their mean and
            # the variance of that mean
            mean_s = a_s["reps"].mean(0)
            var_s = a_s["reps"].var(0, ddof=1)
            ms = a_s["reps"].shape[0]
        else:
            mean_s = a_s["s"] / mf
            # the variance of mean_s: that of one surrogate value (or of one
            # antithetic pair's mean) over the number of values (or pairs)
            ms = mf // 2 if antithetic else mf
            var_s = np.maximum(a_s["s2"] / ms - mean_s ** 2, 0.0)
        mean_d = a_d["d"] / nf
        var_d = np.maximum(a_d["d2"] / nf - mean_d ** 2, 0.0)
        sum_s += w * mean_s
        sum_d += w * mean_d
        var_cv += w * w * (var_d / nf + var_s / ms)
    del nets
    t0 = time.perf_counter()
    p_plain, var_plain = plain_estimate(marks, tscale)
    t["plain"] = time.perf_counter() - t0
    return ControlVariateResult(p_cv=sum_s + sum_d, var_cv=var_cv, p_plain=p_plain,
                                var_plain=var_plain, window=window, times=t, folds=folds,
                                n_traced=n_traced, n_surrogate=n_surrogate,
                                n_surrogate_used=sum(o[3] for o in out), losses=losses,
                                trace_stats=st)


def _node_image(a, window):
    """[..., cells, cell^2] per-cell node values -> [..., nrows * cell, ncols * cell]
    (node k = i * cell + j of cell (r, c) is window node (r * cell + i, c * cell + j))."""
    c, nr, nc = window.cell, window.nrows, window.ncols
    lead = a.shape[:-2]
    a = a.reshape(lead + (nr, nc, c, c))
    a = np.moveaxis(a, -2, -3)                      # [..., nr, c(i), nc, c(j)]
    return a.reshape(lead + (nr * c, nc * c))


def _window_slices(window, shape):
    """(grid rows, grid cols, window rows, window cols) slices of the part of
    the window that lies on a grid of `shape`."""
    h, w = window.nrows * window.cell, window.ncols * window.cell
    r0, c0 = window.row0, window.col0
    gr0, gc0 = max(r0, 0), max(c0, 0)
    gr1, gc1 = min(r0 + h, shape[0]), min(c0 + w, shape[1])
    if gr1 <= gr0 or gc1 <= gc0:
        return None
    return (slice(gr0, gr1), slice(gc0, gc1), slice(gr0 - r0, gr1 - r0),
            slice(gc0 - c0, gc1 - c0))


# The per-node controls: the 3 x 3 cells around a node's own cell (neighbour
# m = cell (r + m // 3 - 1, c + m % 3 - 1)), and the 13 half-offsets of a
# 5 x 5 block whose pair sums give their products (oka_cv_node_sums).
NBR = [(m // 3 - 1, m % 3 - 1) for m in range(9)]
PAIRS = [(0, 0), (0, 1), (0, 2)] + [(1, dc) for dc in range(-2, 3)] + \
        [(2, dc) for dc in range(-2, 3)]


def neighbours(window):
    """[cells, 9] index of each cell's neighbour m (-1 outside the window)."""
    nr, nc = window.nrows, window.ncols
    r, c = np.divmod(np.arange(nr * nc), nc)
    out = np.full((nr * nc, 9), -1, np.int64)
    for m, (dr, dc) in enumerate(NBR):
        rr, cc = r + dr, c + dc
        ok = (rr >= 0) & (rr < nr) & (cc >= 0) & (cc < nc)
        out[ok, m] = rr[ok] * nc + cc[ok]
    return out


def _grid_shape(nb):
    """(rows, columns) of the window's cells from its neighbour table."""
    nc = int(nb[0, 7]) if nb[0, 7] >= 0 else nb.shape[0]   # neighbour (+1, 0) of cell 0
    return nb.shape[0] // nc, nc


def _pair_products(ss, nb):
    """E-sums of s_b s_b' for every pair (m, l) of each cell's neighbours, from
    the half-offset pair sums ss [..., cells, 13]: [..., cells, 9, 9]."""
    hidx = {d: h for h, d in enumerate(PAIRS)}
    out = np.zeros(ss.shape[:-2] + (nb.shape[0], 9, 9))
    for m, (r1, c1) in enumerate(NBR):
        for l, (r2, c2) in enumerate(NBR):
            d = (r2 - r1, c2 - c1)
            if d in hidx:
                base, h = nb[:, m], hidx[d]
            else:
                base, h = nb[:, l], hidx[(-d[0], -d[1])]
            ok = (nb[:, m] >= 0) & (nb[:, l] >= 0)
            out[..., ok, m, l] = ss[..., base[ok], h]
    return out


def node_sums(net, site, d, coef, X, window, parts=2, precision=_lib.FP32, counts=None):
    """oka_cv_node_sums: per part, the sums of s per cell [parts, cells], of
    s_a s_b over the 13 half-offsets [parts, cells, 13], of y per node and of
    y s_b over the node's 9 neighbouring cells [parts, cells, cell^2(, 9)];
    counts (int32 grid, or None) accumulates the realizations' visited
    nodes."""
    pf = site.fixed_field()
    g = pf.grid()
    wxy, s = _site_struct(site)
    starts = start_points(site.xt, site.yt, site.rt, site.npaths)
    o = _lib.make_trace_opts(site.npaths, site.duration, site.tol, site.maxstep, site.umbra,
                             _lib.ptr(starts), precision, site.confined)
    R = len(d["K"])
    ncell, nk = window.nrows * window.ncols, window.cell * window.cell
    out = {"s": np.zeros((parts, ncell)), "ss": np.zeros((parts, ncell, len(PAIRS))),
           "y": np.zeros((parts, ncell, nk)), "ys": np.zeros((parts, ncell, nk, 9))}
    K, n, b, q = (np.ascontiguousarray(d[k], dtype=np.float64) for k in ("K", "n", "b", "q"))
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float32)
    if counts is not None and (counts.dtype != np.int32 or not counts.flags.c_contiguous or
                               counts.shape != (pf.nrows, pf.ncols)):
        raise ValueError("counts must be a C-contiguous int32 array of the grid's shape")
    st = _lib.TraceStats_()
    net.ctx.check(net.ctx.lib.oka_cv_node_sums(
        net.ctx.h, net.h, s, R, _lib.ptr(K), _lib.ptr(n), _lib.ptr(b), _lib.ptr(q),
        _lib.ptr(coef), _lib.ptr(X), o, g, window, int(parts), _lib.ptr(counts),
        _lib.ptr(out["s"]), _lib.ptr(out["ss"]), _lib.ptr(out["y"]), _lib.ptr(out["ys"]), st, 0))
    return out, st


class NodeSet:
    """A traced correction set's node masks, kept on the device
    (oka_cv_node_trace); close() frees them."""

    def __init__(self, ctx, h, nreal, window):
        self.ctx, self.h, self.nreal, self.window = ctx, h, nreal, window

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.oka_node_set_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def node_trace(site, d, coef, window, precision=_lib.FP32, counts=None, ctx=None):
    """Traces the realizations d (coef: their A..F) and keeps their window
    node masks on the device: (NodeSet, trace statistics).  counts (int32
    grid, or None) accumulates their visited nodes."""
    ctx = ctx or _lib.context()
    pf = site.fixed_field()
    g = pf.grid()
    wxy, s = _site_struct(site)
    starts = start_points(site.xt, site.yt, site.rt, site.npaths)
    o = _lib.make_trace_opts(site.npaths, site.duration, site.tol, site.maxstep, site.umbra,
                             _lib.ptr(starts), precision, site.confined)
    K, n, b, q = (np.ascontiguousarray(d[k], dtype=np.float64) for k in ("K", "n", "b", "q"))
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    if counts is not None and (counts.dtype != np.int32 or not counts.flags.c_contiguous or
                               counts.shape != (pf.nrows, pf.ncols)):
        raise ValueError("counts must be a C-contiguous int32 array of the grid's shape")
    st = _lib.TraceStats_()
    h = C.c_void_p()
    ctx.check(ctx.lib.oka_cv_node_trace(ctx.h, s, len(K), _lib.ptr(K), _lib.ptr(n), _lib.ptr(b),
                                        _lib.ptr(q), _lib.ptr(coef), o, g, window,
                                        _lib.ptr(counts), C.byref(h), st, 0))
    return NodeSet(ctx, h, len(K), window), st


def node_reduce(net, nodeset, X, parts=2):
    """The node sums of a NodeSet under the network `net` (oka_cv_node_reduce),
    as node_sums returns them."""
    w = nodeset.window
    ncell, nk = w.nrows * w.ncols, w.cell * w.cell
    out = {"s": np.zeros((parts, ncell)), "ss": np.zeros((parts, ncell, len(PAIRS))),
           "y": np.zeros((parts, ncell, nk)), "ys": np.zeros((parts, ncell, nk, 9))}
    X = np.ascontiguousarray(X, dtype=np.float32)
    if X.shape[0] != nodeset.nreal:
        raise ValueError("one input row per traced realization")
    net.ctx.check(net.ctx.lib.oka_cv_node_reduce(
        net.h, nodeset.h, _lib.ptr(X), int(parts), _lib.ptr(out["s"]), _lib.ptr(out["ss"]),
        _lib.ptr(out["y"]), _lib.ptr(out["ys"]), 0))
    return out


def node_estimate(sums, n_rows, nb, mu_reps, ridge=1e-2, flat=1e-10, min_events=30):
    """The cross-fitted node estimate from node_sums over parts of n_rows[p]
    realizations.  Node j (node k of cell a) is corrected by the surrogate
    outputs s of the cells around a (nb: neighbours(window)):
        p_j = sum_p w_p (ybar_p - beta_(-p) . (sbar_p - mu)),   w_p = n_p / N,
    beta_(-p) the ridge regression of y_j on the controls over the other
    parts' realizations, in correlation form: each control divided by its
    standard deviation, `ridge` added to the correlation matrix's diagonal,
    and controls whose variance is below `flat` (a cell whose output hardly
    moves) left out; a node visited (or missed) by fewer than min_events of
    the other parts' realizations takes no control at all (too few events
    to fit 9 coefficients).  Each part's coefficients depend only on the
    other parts, so the estimate is unbiased whatever the coefficients.  mu_reps [R, cells] are independent unbiased
    replicate means of the surrogate (their mean is mu).  The variance is
        sum_p w_p^2 var_p(y - beta_(-p) . s) / n_p + var(beta_bar . mu),
    beta_bar = sum_p w_p beta_(-p), the last term from the replicates.
    Returns (p, var, beta_bar) as [cells, cell^2] (and [..., 9]) arrays."""
    n_rows = np.asarray(n_rows, dtype=np.float64)
    N = n_rows.sum()
    P = len(n_rows)
    absent = nb < 0
    nbi = np.where(absent, 0, nb)
    mu_reps = np.asarray(mu_reps, dtype=np.float64)
    mu = mu_reps.mean(0)
    tot = {k: v.sum(0) for k, v in sums.items()}

    def moments(S, SS, Y, YS, n):
        """Means, control covariance [cells, 9, 9] and cross-covariance
        [cells, nk, 9] of one set of sums over n realizations."""
        sb = np.where(absent, 0.0, (S / n)[nbi])                   # [cells, 9]
        G = _pair_products(SS, nb) / n - sb[:, :, None] * sb[:, None, :]
        yb = Y / n                                                   # [cells, nk]
        C = YS / n - yb[:, :, None] * sb[:, None, :]
        C = np.where(absent[:, None, :], 0.0, C)
        return sb, G, yb, C

    p = np.zeros(sums["y"].shape[1:])
    var = np.zeros_like(p)
    bbar = np.zeros(p.shape + (9,))
    mu_nb = np.where(absent, 0.0, mu[nbi])
    for h in range(P):
        nh, no = n_rows[h], N - n_rows[h]
        _, Go, _, Co = moments(tot["s"] - sums["s"][h], tot["ss"] - sums["ss"][h],
                               tot["y"] - sums["y"][h], tot["ys"] - sums["ys"][h], no)
        diag = np.einsum("cii->ci", Go)
        off = absent | ~(diag > flat)                                # controls left out
        sd = np.where(off, 1.0, np.sqrt(np.where(off, 1.0, diag)))
        A = Go / (sd[:, :, None] * sd[:, None, :])
        A = np.where(off[:, :, None] | off[:, None, :], 0.0, A)
        A = A + np.eye(9)[None] * np.where(off, 1.0, ridge)[:, :, None]
        rhs = np.where(off[:, None, :], 0.0, Co / sd[:, None, :])   # [cells, nk, 9]
        beta = np.linalg.solve(A, np.swapaxes(rhs, 1, 2))            # [cells, 9, nk]
        beta = np.swapaxes(beta, 1, 2) / sd[:, None, :]              # [cells, nk, 9]
        ev = tot["y"] - sums["y"][h]                                 # the other parts' visits
        beta = np.where((np.minimum(ev, no - ev) < min_events)[:, :, None], 0.0, beta)
        sh, Gh, yh, Ch = moments(sums["s"][h], sums["ss"][h], sums["y"][h], sums["ys"][h], nh)
        w = nh / N
        p += w * (yh - np.matmul(beta, (sh - mu_nb)[:, :, None])[:, :, 0])
        vd = (yh * (1.0 - yh) - 2.0 * (beta * Ch).sum(-1)
              + (np.matmul(beta, Gh) * beta).sum(-1))
        var += w * w * np.maximum(vd, 0.0) / nh
        bbar += w * beta
    # the surrogate term's variance from the replicates' spread: their
    # covariance over each cell's neighbours, through the same pair offsets
    R = mu_reps.shape[0]
    dev = mu_reps - mu                                               # [R, cells]
    nr, nc = _grid_shape(nb)
    dq = np.zeros((nb.shape[0], len(PAIRS)))
    dv = dev.reshape(R, nr, nc)
    for h, (dr, dc) in enumerate(PAIRS):
        if dr >= nr or abs(dc) >= nc:
            continue                     # no pair of the window's cells this far apart
        c0, c1 = max(0, -dc), nc - max(0, dc)
        prod = dv[:, :nr - dr, c0:c1] * dv[:, dr:, c0 + dc:c1 + dc]
        dq.reshape(nr, nc, -1)[:nr - dr, c0:c1, h] = prod.sum(0)
    cov_mu = _pair_products(dq, nb) / ((R - 1) * R)                 # [cells, 9, 9]
    var += (np.matmul(bbar, cov_mu) * bbar).sum(-1)
    return p, var, bbar


class NodeEstimate:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def control_variate_nodes(site, n_train, n_correct, n_surrogate, train_steps=1200, batch=2048,
                          lr=5e-3, seed=0, precision=_lib.FP32, pilot=2000, window=None,
                          n_hidden=512, qmc_replicates=64, qmc_pow2=True, parts=2, overlap=True):
    """Control-variate estimate of every node's capture probability: the
    ProbabilityField of create_stochastic_capturezone (stochastic.py:76-270)
    on the site's fixed lattice.

    n_train traced realizations train one surrogate network on their window
    cells; n_correct further traced realizations and n_surrogate fresh
    surrogate evaluations (qmc_replicates scrambled Sobol' replicates) form
    the estimate of each window node (node_estimate: controls from the 3 x 3
    cells around the node's cell, coefficients cross-fitted over `parts`
    parts of the correction realizations).  Nodes outside the
    window take the plain estimate over all n_train + n_correct traced
    realizations, which are independent draws of the same distribution.
    With overlap, the network trains and sums the surrogate term on a
    second context while the correction set is traced; the node sums follow
    (oka_cv_node_trace, oka_cv_node_reduce).

    Returns a NodeEstimate: p (float64 [nrows, ncols]), var (its variance per
    node: the estimator's in the window, p (1 - p) / n_traced outside), the
    plain estimate p_plain and its variance from the same n_train + n_correct
    realizations, the window, the probability mass outside the window
    (outside_mass, and outside_max its largest node value), and per-phase
    times in seconds."""
    if n_train < batch or n_correct < 2 * parts:
        raise ValueError("too few realizations")
    t = {}
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    R = n_train + n_correct
    d = draw_realizations(site.wells, site.c_dist, site.p_dist, site.t_dist, R, rng)
    wxy = site.wells_xy()
    _, _, coef = fit_and_draw(wxy, site.xt, site.yt, site.base, site.observations, d["K"],
                              d["b"], d["q"], d["z"], moments=False)
    dt = {k: v[:n_train] for k, v in d.items()}
    dc = {k: v[n_train:] for k, v in d.items()}
    t["fit_draw"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    pf = site.fixed_field()
    if window is None:
        npil = min(pilot, n_train)
        pc, _ = trace_counts(site, {k: v[:npil] for k, v in dt.items()}, coef[:npil], precision)
        window = choose_window(pc)
        if window.cell > 4:
            raise ValueError(f"the capture zone needs cells of {window.cell} nodes; node "
                             f"masks take at most 4")
    t["pilot"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    counts = np.zeros((pf.nrows, pf.ncols), np.int32)
    marks, st_train = trace_marks(site, dt, coef[:n_train], window, precision, counts=counts)
    t["trace_train"] = time.perf_counter() - t0
    tscale = 1.0 / (window.cell * window.cell)
    F = features(d, coef)
    std = Standardizer(F[:n_train])
    X = std(F)
    ncell = window.nrows * window.ncols
    # the network on a context of its own: it trains (and sums the surrogate
    # term) while this thread traces the correction set on the default one
    ctx2 = _lib.Context(_lib.context().device)
    net = SurrogateMLP(X.shape[1], n_hidden, ncell, seed=seed + 1, ctx=ctx2)
    npts, nrep = qmc_split(n_surrogate, qmc_replicates, qmc_pow2)
    surr_seed = int(rng.integers(1, 2 ** 63))
    side = {}

    def learn():
        try:
            t1 = time.perf_counter()
            epochs = max(1, int(math.ceil(train_steps / max(1, n_train // batch))))
            side["losses"] = net.train(X[:n_train], marks, tscale, epochs, batch, lr,
                                       seed=seed + 100)
            t2 = time.perf_counter()
            seeds = [_replicate_seed(surr_seed, 0, k) for k in range(nrep)]
            side["reps"] = _surrogate_sums_qmc(net, site, npts, seeds, std) / npts
            side["times"] = (t2 - t1, time.perf_counter() - t2)
        except BaseException as e:   # re-raised in the calling thread
            side["error"] = e

    t0 = time.perf_counter()
    if overlap:
        th = threading.Thread(target=learn)
        th.start()
    else:
        learn()
    t1 = time.perf_counter()
    try:
        nodeset, st_corr = node_trace(site, dc, coef[n_train:], window, precision, counts)
    finally:
        if overlap:
            th.join()
    t["trace_correct"] = time.perf_counter() - t1
    if "error" in side:
        nodeset.close()
        raise side["error"]
    t["train"], t["surrogate"] = side["times"]
    t["train_and_correct"] = time.perf_counter() - t0
    del marks
    t0 = time.perf_counter()
    sums = node_reduce(net, nodeset, X[n_train:], parts)
    nodeset.close()
    reps, losses = side["reps"], side["losses"]
    t["reduce"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    bounds = [n_correct * p // parts for p in range(parts + 1)]
    n_rows = np.diff(bounds)
    p_cells, var_cells, beta = node_estimate(sums, n_rows, neighbours(window), reps)
    p_win, var_win = _node_image(p_cells, window), _node_image(var_cells, window)
    p_plain = counts / float(R)
    var_plain = p_plain * (1.0 - p_plain) / R
    p = p_plain.copy()
    var = var_plain.copy()
    inside = np.zeros(p.shape, bool)
    sl = _window_slices(window, p.shape)
    if sl is not None:
        gr, gc, wr, wc = sl
        p[gr, gc] = p_win[wr, wc]
        var[gr, gc] = var_win[wr, wc]
        inside[gr, gc] = True
    t["estimate"] = time.perf_counter() - t0
    return NodeEstimate(p=p, var=var, p_plain=p_plain, var_plain=var_plain, window=window,
                        inside=inside, outside_mass=float(p_plain[~inside].sum()),
                        outside_max=float(p_plain[~inside].max(initial=0.0)),
                        total_mass=float(p_plain.sum()), n_train=n_train, n_correct=n_correct,
                        n_surrogate=npts * nrep, qmc=(npts, nrep), parts=parts, beta=beta,
                        times=t, losses=losses, trace_stats=(st_train, st_corr))
