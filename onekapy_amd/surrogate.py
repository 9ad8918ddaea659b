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


def trace_marks(site, d, coef, window, precision=_lib.FP32):
    """Per-realization visited-node counts per window cell (uint8 [R, cells])."""
    pf = site.fixed_field()
    g = pf.grid()
    wxy, s = _site_struct(site)
    starts = start_points(site.xt, site.yt, site.rt, site.npaths)
    o = _lib.make_trace_opts(site.npaths, site.duration, site.tol, site.maxstep, site.umbra,
                             _lib.ptr(starts), precision, site.confined)
    R = len(d["K"])
    marks = np.zeros((R, window.nrows * window.ncols), np.uint8)
    st = _lib.TraceStats_()
    K, n, b, q = (np.ascontiguousarray(d[k]) for k in ("K", "n", "b", "q"))
    coef = np.ascontiguousarray(coef)
    ctx = _lib.context()
    ctx.check(ctx.lib.oka_trace_marks(ctx.h, s, R, _lib.ptr(K), _lib.ptr(n), _lib.ptr(b),
                                      _lib.ptr(q), _lib.ptr(coef), o, g, window, _lib.ptr(marks),
                                      None, st, 0))
    return marks, st


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
