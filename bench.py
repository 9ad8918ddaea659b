"""Benchmark of the capture-zone hot path on one MI355X.

One step = one pass of the hot path over one batch of realizations of the
single-site workload: fit + MVN draw of A..F for every realization, then
backtrace + chronicle of every particle into the int32 count grid.  Inputs
(realization parameters, site geometry, start points) are resident in HBM
before the timed region; counts stay on the device.

Workload (default): 25 wells, 40 piezometers, 10^4 realizations x 1000
particles, 20-year backtrace, 400 x 400 nodes at 20 m, umbra 40 m, tol 1 m,
maxstep 10 m.  Synthetic, seeded.

The headline ``value`` is realizations/s of the FP64-exact tracer (the
reference's own arithmetic, bit-exact against the oracle).  The FP32 tracer's
rate is reported beside it (``fp32_gated``) with its capture-probability error
against the FP64-exact tracer on the same realizations (max over nodes, gate
0.01), as are the learned control variate on the same site, the 64-site uneven
batch (configuration 2) and the 200-well stress site with all of its 10^6
realizations (configuration 4), each with its own error sample against
FP64-exact.  The
dominant kernel's roofline fraction and the CPU baselines are in the line too.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}   # MI355X vector peaks


def step_flops(nwells):
    """Canonical flops of one attempted Dormand-Prince step (6 confined
    velocity evaluations of 12 + 10 Nw flops, plus 135 for the stages, error
    estimate, step length and step control)."""
    return 207 + 60 * nwells


def kernel_flops(stats, nwells):
    return (stats.attempted * step_flops(nwells) + 10 * stats.accepted
            + 14 * stats.candidates)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nreal", type=int, default=10000)
    ap.add_argument("--config", default="cfg1", choices=["cfg0", "cfg1", "cfg2", "cfg4"],
                    help="cfg1 (default): the single-site headline workload; cfg2: 64 uneven "
                         "sites in one graph-captured pool launch; cfg4: the 200-well "
                         "2000^2 stress site (--cfg4-nreal realizations)")
    ap.add_argument("--cfg2-scale", type=float, default=1.0,
                    help="cfg2: fraction of each site's realizations")
    ap.add_argument("--cfg4-nreal", type=int, default=1000000,
                    help="cfg4: realizations traced in one launch (the workload's 10^6; each "
                         "launch ends with a tail of about half a second in which the longest "
                         "traces finish, so small subsets understate the workload's rate)")
    ap.add_argument("--cfg4-warmup-nreal", type=int, default=20000,
                    help="cfg4: realizations of each warm-up launch (the first of the same "
                         "draws)")
    ap.add_argument("--knobs", default=None,
                    help="cfg2/cfg4: pool launch shape SLOTS,WAVES_PER_CU,CHUNK instead of "
                         "autotuning")
    ap.add_argument("--cfg4-tune-nreal", type=int, default=20000,
                    help="cfg4: realizations of the autotuning runs (runs much smaller than "
                         "the timed one are bound by their single longest trace and mislead "
                         "the tuner)")
    ap.add_argument("--far-order", type=int, default=0,
                    help="oka_trace_opts.far_order of the FP32 pool legs (cfg2, cfg4): 0 lets "
                         "the library choose, < 0 sums every well directly")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU time budget of each CPU baseline")
    ap.add_argument("--modes", default="fp64,fp32",
                    help="tracers to run (the error gate needs both)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (to scale HBM counters)")
    ap.add_argument("--no-cv", action="store_true", help="skip the control-variate leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the configuration-2 and configuration-4 legs of the default run")
    ap.add_argument("--cfg2-knobs", default="2,16,128",
                    help="pool shape (slots,waves per CU,particles per chunk) of the default run's "
                         "cfg2 leg (the fastest of a sweep on the 64-site batch)")
    ap.add_argument("--cfg4-knobs", default="2,16,64",
                    help="pool shape of the default run's cfg4 leg (the fastest of a sweep on "
                         "100,000 of its realizations)")
    ap.add_argument("--unconfined", action="store_true",
                    help="trace the unconfined velocity (compute_velocity) in the fp64/fp32 modes; "
                         "the roofline keeps the confined flop count")
    ap.add_argument("--only-cv", action="store_true", help="run only the control-variate leg")
    ap.add_argument("--mlp-step", action="store_true",
                    help="time the surrogate network alone: Adam steps of --cv-batch on "
                         "synthetic inputs and inference of one batch (for kernel profiles)")
    ap.add_argument("--cv-train", type=int, default=100000,
                    help="traced realizations the larger estimator's network trains on")
    ap.add_argument("--cv-correct", type=int, default=200000,
                    help="traced realizations that correct the larger estimator")
    ap.add_argument("--cv-surrogate", type=int, default=2000000)
    ap.add_argument("--cv-qmc", type=int, default=64,
                    help="scrambled Sobol' replicates of the surrogate term (a power of two "
                         "points each, so the count rises to fill the budget)")
    ap.add_argument("--cv-reference", type=int, default=1000000,
                    help="realizations of the FP64 plain run the estimates are compared with")
    ap.add_argument("--cv-steps", type=int, default=1200)
    ap.add_argument("--cv-batch", type=int, default=2048)
    ap.add_argument("--cv-lr", type=float, default=5e-3)
    return ap.parse_args()


def control_variate_leg(site, args, fp32_rate):
    """Learned control variate on the same site at node resolution
    (surrogate.control_variate_nodes): T traced realizations (FP32) train the
    network, N further ones correct it node by node with coefficients on the
    3 x 3 cells around each node, plus M surrogate evaluations; nodes outside
    the window take the plain estimate of all T + N.  Both estimators are
    compared node by node with one FP64-exact plain run: the workload's own
    sizes (T = 100,000, as configured, N = 100,000, M = 10^6) and a larger one
    with less error (--cv-train, --cv-correct, --cv-surrogate).  Effective realizations = the
    plain-run size with the same largest node variance, max p (1 - p) /
    max var; per second of the whole pipeline's wall time (draws, fits,
    pilot, tracing, training, surrogate, correction)."""
    from onekapy_amd import _lib
    from onekapy_amd.stochastic import draw_realizations, fit_and_draw
    from onekapy_amd.surrogate import control_variate_nodes, trace_counts

    def run(n_train, n_correct, n_surrogate):
        print(f"bench: control variate: {n_train} traced to train, {n_correct} to correct, "
              f"{n_surrogate} surrogate", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        res = control_variate_nodes(site, n_train, n_correct, n_surrogate,
                                    train_steps=args.cv_steps, batch=args.cv_batch,
                                    lr=args.cv_lr, seed=123, precision=_lib.FP32,
                                    qmc_replicates=args.cv_qmc)
        wall = time.perf_counter() - t0
        print(f"bench: control variate {wall:.2f} s {res.times}", file=sys.stderr, flush=True)
        return res, wall

    # untimed warm-up of the code paths the timed runs take (first launches,
    # first use of buffers), as the tracer legs' warmup steps do
    print("bench: control variate warm-up", file=sys.stderr, flush=True)
    control_variate_nodes(site, 8192, 8192, 1 << 17, train_steps=64, batch=args.cv_batch,
                          lr=args.cv_lr, seed=7, precision=_lib.FP32, qmc_replicates=16)
    tuned = run(args.cv_train, args.cv_correct, args.cv_surrogate)
    configured = run(100000, 100000, 1000000)
    rng = np.random.default_rng(321)
    d = draw_realizations(site.wells, site.c_dist, site.p_dist, site.t_dist, args.cv_reference, rng)
    _, _, coef = fit_and_draw(site.wells_xy(), site.xt, site.yt, site.base, site.observations,
                              d["K"], d["b"], d["q"], d["z"])
    print(f"bench: FP64 reference run of {args.cv_reference} realizations", file=sys.stderr, flush=True)
    tj = time.perf_counter()
    counts, _ = trace_counts(site, d, coef, _lib.FP64_EXACT)
    tj = time.perf_counter() - tj
    p_ref = counts / float(args.cv_reference)
    var_ref = p_ref * (1 - p_ref) / args.cv_reference

    def summary(res, wall):
        n_tr = res.n_train + res.n_correct
        err = np.abs(res.p - p_ref)
        err_plain = np.abs(res.p_plain - p_ref)
        pc = np.clip(res.p, 0.0, 1.0)
        n_eff = (pc * (1 - pc)).max() / res.var.max()
        live = (p_ref > 0.02) & (p_ref < 0.98)
        z = (res.p - p_ref)[live] / np.sqrt(res.var + var_ref)[live]
        eff_rate = n_eff / wall
        w = res.window
        return {
            "traced_train": res.n_train, "traced_correct": res.n_correct,
            "surrogate_evals": res.n_surrogate,
            "surrogate_sampling": f"scrambled Sobol' points, {res.qmc[1]} replicates of "
                                  f"{res.qmc[0]}",
            "window": {"row0": w.row0, "col0": w.col0, "cell_nodes": w.cell,
                       "cells": [w.nrows, w.ncols]},
            "nodes": int(res.p.size),
            "max_error": float(err.max()), "l1_error": float(err.sum()),
            "max_error_plain_same_traced": float(err_plain.max()),
            "l1_error_plain_same_traced": float(err_plain.sum()),
            "rms_z_vs_reference": float(np.sqrt(np.mean(z ** 2))),
            "max_std": float(np.sqrt(res.var.max())),
            "max_std_plain_same_traced": float(np.sqrt(res.var_plain.max())),
            "variance_ratio_at_max": float(res.var_plain.max() / res.var.max()),
            "outside_window_mass": res.outside_mass, "total_mass": res.total_mass,
            "outside_window_max": res.outside_max,
            "effective_realizations": float(n_eff),
            "effective_per_traced": float(n_eff / n_tr),
            "wall_seconds": wall, "phase_seconds": res.times,
            "effective_realizations_per_s": float(eff_rate),
            "vs_fp32_plain_rate": float(eff_rate / fp32_rate) if fp32_rate else None,
        }

    out = summary(*tuned)
    out.update({"resolution": "nodes", "train_steps": args.cv_steps, "batch": args.cv_batch,
                "lr": args.cv_lr, "network": "34->512->512->4096 GELU, BF16 MFMA",
                "controls": "the 3 x 3 cells around each node, cross-fitted over 2 parts",
                "reference_run": {"realizations": args.cv_reference, "tracer": "fp64",
                                  "seconds": tj}})
    out["as_configured"] = summary(*configured)
    return out


def mlp_step(args):
    """The control variate's network alone (34 -> 512 -> 512 -> 4096): Adam
    steps of --cv-batch samples on synthetic inputs and window marks, through
    oka_mlp_train with the sample orders drawn beforehand (the host's shuffle
    is not timed), one untimed call first."""
    from onekapy_amd.surrogate import SurrogateMLP
    rng = np.random.default_rng(0)
    b = args.cv_batch
    n = 16 * b
    x = rng.normal(0, 1, (n, 34)).astype(np.float32)
    marks = rng.integers(0, 17, (n, 4096)).astype(np.uint8)
    net = SurrogateMLP(34, 512, 4096, max_batch=b, seed=1)
    lib = net.ctx.lib

    def train(steps):
        order = rng.integers(0, n, steps * b).astype(np.int32)
        t0 = time.perf_counter()
        net.ctx.check(lib.oka_mlp_train(net.h, n, x.ctypes.data, marks.ctypes.data, 1.0 / 16,
                                        order.ctypes.data, b, steps, args.cv_lr, 0.9, 0.999,
                                        1e-8, None, 0))
        return time.perf_counter() - t0

    train(16)
    wall = {32: [], 160: []}
    for _ in range(3):
        for steps in wall:
            wall[steps].append(train(steps))
    wall = {k: min(v) for k, v in wall.items()}
    # every call stages the same samples once: the difference is 128 steps
    step = (wall[160] - wall[32]) / 128
    flops = 2.0 * b * (34 * 512 + 512 * 512 + 512 * 4096)
    return {"network": "34->512->512->4096 GELU, BF16 MFMA", "batch": b,
            "train_step_us": step * 1e6, "train_tflops": 3 * flops / step / 1e12,
            "call_seconds": {"32 steps": wall[32], "160 steps": wall[160]},
            "note": "per step: the difference of the fastest of 3 calls of 160 and of 32 "
                    "steps, over 128"}


def copy_calibration(torch):
    """A plain copy of 2^28 float32 (1 GiB read + 1 GiB write) for scaling the
    FETCH_SIZE / WRITE_SIZE counters."""
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30}))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_traffic_cfg1.json")


def measured_traffic(kernel_prefix, nreal):
    """HBM bytes per launch of the named kernel at the default workload, from
    the committed rocprofv3 counter summary (None when absent or not
    applicable to this batch size)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except OSError:
        return None
    k = t.get("kernels", {}).get(kernel_prefix)
    if not k or t.get("realizations") != nreal:
        return None
    return k["bytes_per_launch"]


class DeviceRun:
    """Inputs of one site resident on the GPU, and the launches of one step."""

    def __init__(self, site, d, torch, confined=True):
        from onekapy_amd import _lib
        from onekapy_amd.capturezone import start_points
        self.torch = torch
        self.lib = _lib
        self.ctx = _lib.context(0)
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        dev = torch.device("cuda:0")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
        self.site = site
        self.confined = confined
        self.R = len(d["K"])
        self.K, self.n, self.b = t(d["K"]), t(d["n"]), t(d["b"])
        self.q, self.z = t(d["q"]), t(d["z"])
        self.coef = torch.zeros((self.R, 6), dtype=torch.float64, device=dev)
        self.wells = t(site.wells_xy())
        self.obs = t(site.obs_array())
        starts = start_points(site.xt, site.yt, site.rt, site.npaths)
        self.starts = t(starts)
        pf = site.fixed_field()
        self.grid = _lib.make_grid(pf.xmin, pf.ymin, pf.xmax, pf.ymax, pf.deltax, pf.deltay,
                                   pf.nrows, pf.ncols)
        self.counts = torch.zeros((pf.nrows, pf.ncols), dtype=torch.int32, device=dev)
        self.status = torch.zeros(self.R, dtype=torch.int32, device=dev)
        self.site_s = _lib.make_site(None, site.xt, site.yt, site.base, None,
                                     wells_ptr=self.wells.data_ptr(), obs_ptr=self.obs.data_ptr())
        self.site_s.nwells = self.wells.shape[0]
        self.site_s.nobs = self.obs.shape[0]
        torch.cuda.synchronize()

    def step(self, precision, **sched):
        """fit + draw, then trace + chronicle; returns the trace launch's stats.
        sched: scheduling options of oka_trace_opts (schedule, pool_*)."""
        L, c = self.lib, self.ctx
        self.counts.zero_()
        c.check(c.lib.oka_fit_draw(c.h, self.site_s, self.R, self.K.data_ptr(), self.b.data_ptr(),
                                   self.q.data_ptr(), self.z.data_ptr(), None, None,
                                   self.coef.data_ptr(), self.status.data_ptr(),
                                   L.PTR_DEVICE | L.ASYNC))
        s = self.site
        opts = L.make_trace_opts(s.npaths, s.duration, s.tol, s.maxstep, s.umbra,
                                 self.starts.data_ptr(), precision, self.confined, **sched)
        stats = L.TraceStats_()
        c.check(c.lib.oka_trace_chronicle(c.h, self.site_s, self.R, self.K.data_ptr(),
                                          self.n.data_ptr(), self.b.data_ptr(), self.q.data_ptr(),
                                          self.coef.data_ptr(), opts, self.grid, 0, None, 1,
                                          self.counts.data_ptr(), None, None, None, stats,
                                          L.PTR_DEVICE))
        return stats


def time_mode(run, precision, steps, warmup, torch, **sched):
    for _ in range(warmup):
        run.step(precision, **sched)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms, st = [], None
    for _ in range(steps):
        st = run.step(precision, **sched)
        kms.append(st.kernel_ms)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    counts = run.counts.cpu().numpy().copy()
    return dict(seconds=dt, ms_per_step=1e3 * dt / steps, kernel_ms=float(np.mean(kms)),
                stats=st, counts=counts)


def pyref_realizations(site, d, reals, budget_s, stride, confined):
    """The numpy-scalar restatement of the reference (oracle/pyref.py) on
    realizations `reals` of draws d: per realization the fit (lstsq + inv),
    the MVN draw, then every stride-th particle traced and rasterized, until
    budget_s of wall time is spent.  Returns (traces, attempted steps,
    seconds)."""
    from oracle import pyref
    from onekapy_amd.sites import fixed_bounds
    t0 = time.perf_counter()
    traces = steps = 0
    wxy = site.wells_xy()
    for r in reals:
        m = pyref.PyModel(site.base, d["K"][r], d["n"][r], d["b"][r],
                          [(wxy[i, 0], wxy[i, 1], 0.25, d["q"][r][i]) for i in range(len(wxy))])
        ev, cov = m.fit(site.obs_array(), site.xt, site.yt)
        m.coef = np.random.default_rng(r).multivariate_normal(ev.reshape(6), cov)   # stochastic.py:241
        pf = pyref.PyField(site.spacing, site.spacing, site.xt, site.yt)
        pf.expand(*fixed_bounds(site.xt, site.yt, site.spacing, site.nrows, site.ncols))
        vel = m.velocity_confined if confined else m.velocity

        def feval(xy):
            vx, vy = vel(xy[0], xy[1])
            return np.array([-vx, -vy])

        for xs, ys in pyref.start_points(site.xt, site.yt, site.rt, site.npaths)[::stride]:
            v, att, _acc, _st = pyref.backtrace(xs, ys, site.duration, site.tol, site.maxstep,
                                                feval)
            pf.rasterize([a[0] for a in v], [a[1] for a in v], site.umbra)
            traces += 1
            steps += att
            if time.perf_counter() - t0 > budget_s:
                return traces, steps, time.perf_counter() - t0
        pf.register(1.0)
    return traces, steps, time.perf_counter() - t0


def _pyref_worker(job):
    return pyref_realizations(*job)


def cpu_baselines(site, d, budget_s, confined=True):
    """Reference-speed CPU baselines on bounded samples of the workload, each
    realization fitted, drawn and traced on the CPU: the numpy-scalar
    restatement in one process (how the reference runs) and in one process per
    host core over disjoint realizations, and the C oracle with OpenMP.  Run
    before the process touches the GPU."""
    import multiprocessing as mp
    import oracle as O
    from onekapy_amd.capturezone import start_points
    out = {}
    P = site.npaths
    # (1) one process: every 4th particle of the first 4 realizations
    tr, steps, el = pyref_realizations(site, d, range(4), budget_s, 4, confined)
    out["python"] = {"value": tr / P / el, "unit": "realizations/s", "cores": 1, "kind": "port",
                     "particle_steps_per_s": steps / el,
                     "sample": f"{tr} traces (every 4th particle of realizations 0-3, each "
                               f"fitted and drawn; {steps} attempted steps, {el:.1f} s), "
                               f"numpy-scalar restatement with the reference's loop structure, "
                               f"one process"}
    # (2) one process per core (at most 16, the box's share), realization i in process i
    nproc = max(1, min(16, os.cpu_count() or 1))
    jobs = [(site, {k: v[:i + 1] for k, v in d.items()}, [i], budget_s, 1, confined)
            for i in range(nproc)]
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(nproc) as pool:
        res = pool.map(_pyref_worker, jobs)
    wall = time.perf_counter() - t0
    tr = sum(r[0] for r in res)
    steps = sum(r[1] for r in res)
    el = max(r[2] for r in res)
    out["python_processes"] = {
        "value": tr / P / el, "unit": "realizations/s", "cores": nproc, "kind": "port",
        "particle_steps_per_s": steps / el,
        "sample": f"{nproc} processes, realization i in process i ({tr} traces, {steps} "
                  f"attempted steps, {el:.1f} s per process, {wall:.1f} s with process start)"}
    # (3) C oracle, FP64, OpenMP over realizations
    threads = O.max_threads()
    pf = site.fixed_field()
    starts = start_points(site.xt, site.yt, site.rt, P)
    g = O.make_grid(pf.xmin, pf.ymin, pf.xmax, pf.ymax, pf.deltax, pf.deltay, pf.nrows, pf.ncols)
    opts = O.trace_opts(site.duration, site.tol, site.maxstep, confined, O.MATH_GLIBC)
    wxy = site.wells_xy()

    def oracle_coef(n):
        coef = np.zeros((n, 6))
        for r in range(n):
            om = O.OracleModel(site.base, d["K"][r], d["n"][r], d["b"][r],
                               np.c_[wxy, np.full(len(wxy), 0.25), d["q"][r]])
            ev, cov = om.fit(site.obs_array(), site.xt, site.yt, O.MATH_GLIBC)
            coef[r] = O.mvn(ev, cov, d["z"][r])
        return coef

    nr = max(threads, 2)
    while True:
        t0 = time.perf_counter()
        coef = oracle_coef(nr)
        _, st = O.chronicle_fixed(wxy, site.xt, site.yt, site.base, d["K"][:nr], d["n"][:nr],
                                  d["b"][:nr], d["q"][:nr], coef, starts, site.umbra, opts, g,
                                  nthreads=threads)
        el = time.perf_counter() - t0
        if el > 0.3 * budget_s or nr >= len(d["K"]):
            break
        nr = min(len(d["K"]), int(nr * max(2.0, 0.5 * budget_s / max(el, 1e-3))))
    out["c_oracle"] = {"value": nr / el, "unit": "realizations/s", "cores": threads,
                       "kind": "port", "particle_steps_per_s": st[0] / el,
                       "sample": f"first {nr} realizations x {P} particles, fits and draws "
                                 f"included ({int(st[0])} attempted steps, {el:.1f} s), C FP64 "
                                 f"oracle, OpenMP"}
    return out


def oracle_sample_rate(site, d, coef, budget_s, max_real=None, confined=True):
    """The C FP64 oracle (OpenMP, all host threads) on the first realizations
    of a site, growing the sample until it takes a sizeable part of budget_s."""
    import oracle as O
    from onekapy_amd.capturezone import start_points
    pf = site.fixed_field()
    starts = start_points(site.xt, site.yt, site.rt, site.npaths)
    g = O.make_grid(pf.xmin, pf.ymin, pf.xmax, pf.ymax, pf.deltax, pf.deltay, pf.nrows, pf.ncols)
    opts = O.trace_opts(site.duration, site.tol, site.maxstep, confined, O.MATH_GLIBC)
    threads = O.max_threads()
    cap = len(d["K"]) if max_real is None else min(max_real, len(d["K"]))
    nr = min(cap, max(1, threads // 4))
    while True:
        t0 = time.perf_counter()
        _, st = O.chronicle_fixed(site.wells_xy(), site.xt, site.yt, site.base, d["K"][:nr],
                                  d["n"][:nr], d["b"][:nr], d["q"][:nr], coef[:nr], starts,
                                  site.umbra, opts, g, nthreads=threads)
        el = time.perf_counter() - t0
        if el > 0.3 * budget_s or nr >= cap:
            return nr, el, int(st[0]), threads
        nr = min(cap, int(nr * max(2.0, 0.5 * budget_s / max(el, 1e-3))))


def fixed_knobs(args):
    """(knobs, []) from --knobs, or (None, None) to autotune."""
    if not args.knobs:
        return None, None
    sl, wpc, ch = (int(v) for v in args.knobs.split(","))
    return {"pool_slots": sl, "pool_waves_per_cu": wpc, "pool_chunk": ch}, []


FLOP_MODEL_FAR = ("canonical: every attempted step priced at the direct sum over all wells "
                  "(207 + 60 per well); on sites of 64 wells or more the far-well expansions "
                  "execute fewer (cfg4: 5.6x fewer lane-instructions per step), so this is a "
                  "direct-sum-equivalent rate")


def canonical_flops(stats, nwells):
    return stats.attempted * (207 + 60 * nwells) + 10 * stats.accepted + 14 * stats.candidates


def bench_cfg2(args):
    """64 uneven sites in one persistent, graph-captured pool launch (autotuned)."""
    from onekapy_amd import _lib, sites
    from onekapy_amd.autotune import tune_batch
    from onekapy_amd.batch import SiteBatch
    from onekapy_amd.surrogate import trace_counts
    log = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
    best, trials = fixed_knobs(args)
    if best is None:
        small = sites.config2(scale=0.05)
        sdraws = [sites.draw(s, seed=100 + i) for i, s in enumerate(small)]
        tb = SiteBatch(small, sdraws, far_order=args.far_order)
        log("cfg2: autotuning the pool launch on a 5 % copy")
        best, trials = tune_batch(tb, log=log)
        tb.close()
    ss = sites.config2(scale=args.cfg2_scale)
    draws = [sites.draw(s, seed=100 + i) for i, s in enumerate(ss)]
    b = SiteBatch(ss, draws, knobs=best, far_order=args.far_order)
    R = sum(b.nreal)
    log(f"cfg2: 64 sites, {R} realizations x 256 particles: {args.warmup} warmup + "
        f"{args.steps} timed replays")
    for _ in range(args.warmup):
        b.run()
    t0 = time.perf_counter()
    kms = [b.run() for _ in range(args.steps)]
    dt = time.perf_counter() - t0
    res = [b.results(i) for i in range(len(ss))]
    att = sum(r[2].attempted for r in res)
    flops = sum(canonical_flops(r[2], len(s.wells)) for r, s in zip(res, ss))
    kernel_ms = float(np.mean(kms))
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    # FP32 (the batch's arithmetic: the pool tracer with the same far-well
    # choice) against the FP64-exact tracer on the first 400 realizations of 6
    # sites (one node flipped in one realization is then 0.0025)
    errs = []
    for i in range(6):
        n = min(400, b.nreal[i])
        dd = {k: v[:n] for k, v in draws[i].items()}
        coef = res[i][1][:n]
        c64, _ = trace_counts(ss[i], dd, coef, _lib.FP64_EXACT)
        c32, _ = trace_counts(ss[i], dd, coef, _lib.FP32, _lib.SCHED_POOL, args.far_order)
        errs.append(np.abs(c32 - c64) / n)
    err_max = float(max(e.max() for e in errs))
    err_l1 = float(sum(e.sum() for e in errs))
    b.close()
    # the same 64 sites in FP64-exact: the pool in the workgroup tracer's arithmetic
    log("cfg2: FP64-exact batch, one warm-up and one timed replay")
    b64 = SiteBatch(ss, draws, precision=_lib.FP64_EXACT)
    b64.run()
    t0 = time.perf_counter()
    ms64 = b64.run()
    dt64 = time.perf_counter() - t0
    res64 = [b64.results(i) for i in range(len(ss))]
    flops64 = sum(canonical_flops(r[2], len(s.wells)) for r, s in zip(res64, ss))
    ach64 = flops64 / (ms64 * 1e-3) / 1e12
    fp64 = {"value": R / dt64, "unit": "realizations/s", "dtype": "f64", "seconds": dt64,
            "particle_steps_per_s": sum(r[2].attempted for r in res64) / dt64,
            "roofline": {"bound": "valu", "achieved": ach64, "peak": PEAK_TFLOPS["fp64"],
                         "unit": "TFLOP/s", "frac": ach64 / PEAK_TFLOPS["fp64"],
                         "kernel": "pool tracer (fp64-exact), whole batch",
                         "flops_per_launch": flops64}}
    b64.close()
    cpu = None
    if not args.no_cpu:
        log("cfg2: CPU baseline (C oracle) on site 0")
        nr, el, st0, threads = oracle_sample_rate(ss[0], draws[0], res[0][1], args.cpu_seconds)
        cpu = {"value": nr / el, "unit": "realizations/s", "cores": threads, "kind": "port",
               "particle_steps_per_s": st0 / el,
               "sample": f"site 0 ({len(ss[0].wells)} wells), first {nr} realizations x 256 "
                         f"particles ({st0} attempted steps, {el:.1f} s), C FP64 oracle, OpenMP"}
    peak = PEAK_TFLOPS["fp32"]
    return {
        "metric": "realizations/sec and particle-steps/sec at max capture-prob error <=0.01 vs FP64",
        "value": R * args.steps / dt, "unit": "realizations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded sites and realization parameters)",
        "config": {"workload": "cfg2: 64 sites, wells U{1..200}, obs U{10..100}, grids "
                               "U{100..800}^2 @ 20 m, realizations U{100..5000} x 256, "
                               "backtraces U(1, 50) y", "sites": len(ss), "realizations": R,
                   "particles": 256, "scale": args.cfg2_scale, "autotuned": best},
        "particle_steps_per_s": att * args.steps / dt,
        "error_vs_fp64": {"max": err_max, "l1": err_l1,
                          "sample": "first 400 realizations (or all) of sites 0-5, FP32 "
                                    "pool tracer as timed against the FP64-exact tracer"},
        "autotune": [{"knobs": k, "ms": ms} for k, ms in trials],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "pool tracer (fp32), whole batch", "flops_per_launch": flops,
                     "flop_model": FLOP_MODEL_FAR},
        "fp64_exact": fp64,
        "cpu_baseline": cpu,
    }


def bench_cfg4(args, torch):
    """The 200-well, 2000^2 stress site (all 10^6 realizations by default),
    FP32 pool tracer."""
    from onekapy_amd import _lib, sites
    from onekapy_amd.autotune import tune_batch
    from onekapy_amd.batch import SiteBatch
    from onekapy_amd.surrogate import trace_counts
    log = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
    site = sites.config4(args.cfg4_nreal)
    d = sites.draw(site, seed=4)
    best, trials = fixed_knobs(args)
    if best is None:
        small = sites.config4(args.cfg4_tune_nreal)
        tb = SiteBatch([small], [sites.draw(small, seed=40)], far_order=args.far_order)
        log(f"cfg4: autotuning the pool launch on {args.cfg4_tune_nreal} realizations")
        # one sweep over the knobs that matter here (each replay takes seconds)
        search = (("pool_waves_per_cu", (8, 12, 16)), ("pool_slots", (2, 3)),
                  ("pool_chunk", (32, 64)))
        best, trials = tune_batch(tb, log=log, search=search, passes=1)
        tb.close()
    sched = dict(schedule=_lib.SCHED_POOL, far_order=args.far_order, **best)
    if args.warmup > 0:
        nw = min(args.cfg4_warmup_nreal, len(d["K"]))
        log(f"cfg4: {args.warmup} warm-up launches of the first {nw} realizations")
        wrun = DeviceRun(site, {k: v[:nw] for k, v in d.items()}, torch)
        for _ in range(args.warmup):
            wrun.step(_lib.FP32, **sched)
        torch.cuda.synchronize()
        del wrun
    run = DeviceRun(site, d, torch)
    log(f"cfg4: {args.cfg4_nreal} realizations: {args.steps} timed steps")
    m = time_mode(run, _lib.FP32, args.steps, 0, torch, **sched)
    R = run.R
    st = m["stats"]
    flops = canonical_flops(st, len(site.wells))
    achieved = flops / (m["kernel_ms"] * 1e-3) / 1e12
    # FP32 against FP64-exact on the first 2000 realizations
    n = min(2000, R)
    coef = run.coef.cpu().numpy()
    dd = {k: v[:n] for k, v in d.items()}
    log(f"cfg4: FP64-exact subset of {n} realizations (pool schedule)")
    t64 = time.perf_counter()
    c64, st64 = trace_counts(site, dd, coef[:n], _lib.FP64_EXACT, _lib.SCHED_POOL)
    t64 = time.perf_counter() - t64
    c32, _ = trace_counts(site, dd, coef[:n], _lib.FP32, _lib.SCHED_POOL, args.far_order)
    err = np.abs(c32 - c64) / n
    cpu = None
    if not args.no_cpu:
        log("cfg4: CPU baseline (C oracle)")
        nr, el, st0, threads = oracle_sample_rate(site, d, coef, args.cpu_seconds, max_real=64)
        cpu = {"value": nr / el, "unit": "realizations/s", "cores": threads, "kind": "port",
               "particle_steps_per_s": st0 / el,
               "sample": f"first {nr} realizations x 256 particles ({st0} attempted steps, "
                         f"{el:.1f} s), C FP64 oracle, OpenMP"}
    peak = PEAK_TFLOPS["fp32"]
    return {
        "metric": "realizations/sec and particle-steps/sec at max capture-prob error <=0.01 vs FP64",
        "value": R * args.steps / m["seconds"], "unit": "realizations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded site and realization parameters)",
        "config": {"workload": site.name, "realizations": R, "of": 1000000, "particles": 256,
                   "wells": len(site.wells), "grid": [site.nrows, site.ncols],
                   "schedule": "pool", "autotuned": best},
        "particle_steps_per_s": st.attempted * args.steps / m["seconds"],
        "status_count": [int(v) for v in st.status_count],
        "error_vs_fp64": {"max": float(err.max()), "l1": float(err.sum()),
                          "sample": f"first {n} realizations, FP32 pool tracer as timed "
                                    f"against the FP64-exact tracer on the pool schedule "
                                    f"({t64:.1f} s)"},
        "autotune": [{"knobs": k, "ms": ms} for k, ms in trials],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "pool tracer (fp32)", "flops_per_launch": flops,
                     "flop_model": FLOP_MODEL_FAR},
        "cpu_baseline": cpu,
    }


def _embedded(result, keys=("value", "unit", "ms_per_step", "particle_steps_per_s", "dtype",
                             "error_vs_fp64", "roofline", "cpu_baseline", "config", "status_count",
                             "steps", "warmup", "fp64_exact")):
    return {k: result[k] for k in keys if k in result}


def cfg2_leg(knobs, args):
    """Configuration 2 inside the default run: one warm-up and one timed replay
    of the 64-site batch with a fixed pool shape, plus its FP64-exact error
    sample and the C oracle on site 0."""
    a = argparse.Namespace(**vars(args))
    a.steps, a.warmup, a.cfg2_scale = 1, 1, 1.0
    a.knobs = ",".join(str(knobs[k]) for k in ("pool_slots", "pool_waves_per_cu", "pool_chunk"))
    a.cpu_seconds = min(args.cpu_seconds, 8.0)
    return _embedded(bench_cfg2(a))


def cfg4_leg(knobs, args, torch):
    """Configuration 4 inside the default run: all 10^6 realizations in one
    timed launch with a fixed pool shape (after a warm-up launch of the first
    20,000), plus its FP64-exact error sample and the C oracle on its first
    realizations."""
    a = argparse.Namespace(**vars(args))
    a.steps, a.warmup, a.cfg4_nreal, a.cfg4_warmup_nreal = 1, 1, 1000000, 20000
    a.knobs = ",".join(str(knobs[k]) for k in ("pool_slots", "pool_waves_per_cu", "pool_chunk"))
    a.cpu_seconds = min(args.cpu_seconds, 8.0)
    return _embedded(bench_cfg4(a, torch))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(realizations shard across GPUs with one int32 sum-reduce of the count grid, "
              f"not built here)", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calibration:
        copy_calibration(torch)
        return
    from onekapy_amd import sites
    from onekapy_amd import _lib

    if args.config == "cfg2":
        print(json.dumps(bench_cfg2(args)))
        return
    if args.config == "cfg4":
        print(json.dumps(bench_cfg4(args, torch)))
        return
    site = sites.config1(args.nreal) if args.config == "cfg1" else sites.config0()
    if args.mlp_step:
        print(json.dumps({"mlp_step": mlp_step(args)}))
        return
    if args.only_cv:
        cv = control_variate_leg(site, args, None)
        print(json.dumps({"control_variate": cv}))
        return
    d = sites.draw(site, seed=1)
    confined = not args.unconfined
    cpu = None
    if not args.no_cpu:
        # before the process touches the GPU: the multi-process leg starts workers
        print("bench: CPU baselines", file=sys.stderr, flush=True)
        cpu = cpu_baselines(site, d, args.cpu_seconds, confined)
    run = DeviceRun(site, d, torch, confined=confined)
    nw = len(site.wells)
    modes = {}
    wanted = [m.strip() for m in args.modes.split(",") if m.strip()]
    for name, prec in (("fp64", _lib.FP64_EXACT), ("fp32", _lib.FP32)):
        if name not in wanted:
            continue
        print(f"bench: {name} {site.name}: {args.warmup} warmup + {args.steps} timed steps",
              file=sys.stderr, flush=True)
        modes[name] = time_mode(run, prec, args.steps, args.warmup, torch)
        m = modes[name]
        print(f"bench: {name} {m['ms_per_step']:.1f} ms/step, trace kernel {m['kernel_ms']:.1f} ms, "
              f"{m['stats'].attempted} steps", file=sys.stderr, flush=True)
    R = run.R
    err_max = err_l1 = None
    if "fp64" in modes and "fp32" in modes:
        p64 = modes["fp64"]["counts"] / R
        p32 = modes["fp32"]["counts"] / R
        err_max = float(np.abs(p32 - p64).max())
        err_l1 = float(np.abs(p32 - p64).sum())
    head = "fp64" if "fp64" in modes else next(iter(modes))
    hm = modes[head]
    value = R * args.steps / hm["seconds"]
    steps_per_s = hm["stats"].attempted * args.steps / hm["seconds"]
    flops = kernel_flops(hm["stats"], nw)
    achieved = flops / (hm["kernel_ms"] * 1e-3) / 1e12
    peak = PEAK_TFLOPS[head]

    cv = None
    if not args.no_cv and "fp32" in modes and confined:
        cv = control_variate_leg(site, args, R * args.steps / modes["fp32"]["seconds"])

    extra = {}
    if not args.no_extra and args.config == "cfg1":
        def knobs(text):
            return dict(zip(("pool_slots", "pool_waves_per_cu", "pool_chunk"),
                            (int(v) for v in text.split(","))))
        extra["config2"] = cfg2_leg(knobs(args.cfg2_knobs), args)
        extra["config4"] = cfg4_leg(knobs(args.cfg4_knobs), args, torch)

    def mode_summary(name):
        m = modes[name]
        return {"realizations_per_s": R * args.steps / m["seconds"],
                "particle_steps_per_s": m["stats"].attempted * args.steps / m["seconds"],
                "ms_per_step": m["ms_per_step"], "trace_kernel_ms": m["kernel_ms"],
                "attempted_steps": int(m["stats"].attempted),
                "accepted_steps": int(m["stats"].accepted),
                "visited_nodes": int(m["stats"].visited),
                "escaped_traces": int(m["stats"].escaped),
                "status_count": [int(v) for v in m["stats"].status_count],
                "kernel_tflops": kernel_flops(m["stats"], nw) / (m["kernel_ms"] * 1e-3) / 1e12}

    fp32_gated = None
    if "fp32" in modes and err_max is not None:
        m32 = modes["fp32"]
        fp32_gated = {"value": R * args.steps / m32["seconds"], "unit": "realizations/s",
                      "dtype": "f32", "gate": "max |p32 - p64| over nodes <= 0.01",
                      "passed": err_max <= 0.01, "error_vs_fp64": {"max": err_max, "l1": err_l1},
                      "particle_steps_per_s": m32["stats"].attempted * args.steps / m32["seconds"],
                      "roofline": {"bound": "valu",
                                   "achieved": mode_summary("fp32")["kernel_tflops"],
                                   "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                                   "frac": mode_summary("fp32")["kernel_tflops"]
                                   / PEAK_TFLOPS["fp32"],
                                   "traffic": measured_traffic("k_trace_fp32", R)}}

    line = {
        "metric": "realizations/sec and particle-steps/sec at max capture-prob error <=0.01 vs FP64",
        "value": value,
        "unit": "realizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32" else "f64",
        "data": "synthetic (seeded site and realization parameters)",
        "config": {"workload": site.name, "realizations": R, "particles": site.npaths,
                   "wells": nw, "observations": len(site.observations),
                   "duration_d": site.duration, "grid": [site.nrows, site.ncols],
                   "spacing_m": site.spacing, "umbra_m": site.umbra, "tol_m": site.tol,
                   "maxstep_m": site.maxstep, "headline_mode": head,
                   "velocity": "confined" if confined else "unconfined"},
        "particle_steps_per_s": steps_per_s,
        "error_vs_fp64": {"max": 0.0 if head == "fp64" else err_max,
                          "l1": 0.0 if head == "fp64" else err_l1,
                          "nodes": int(site.nrows * site.ncols),
                          "note": "the FP64-exact tracer is the reference arithmetic itself"
                                  if head == "fp64" else "FP32 against FP64-exact"},
        "modes": {k: mode_summary(k) for k in modes},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": measured_traffic(f"k_trace_{head}", R),
                     "kernel": f"trace+chronicle ({head})",
                     "flops_per_launch": flops},
        "cpu_baseline": None if cpu is None else dict(cpu["python"],
                                                       processes=cpu["python_processes"],
                                                       c_oracle=cpu["c_oracle"]),
        "fp32_gated": fp32_gated,
        "control_variate": cv,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
