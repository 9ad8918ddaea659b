"""GPU edge cases and size-independent properties, each against the oracle."""
import numpy as np
import pytest

import oracle as O
from onekapy_amd import (Realizations, chronicle_into_field, sites, start_points, trace_vertices,
                         _lib)
from onekapy_amd.sites import fixed_field
from helpers import oracle_chronicle, oracle_fit_draw, oracle_grid, small_site

pytestmark = pytest.mark.gpu


def _one_model(wells, K=10.0, n=0.25, b=20.0, coef=(0, 0, 0, 0, 0, 0)):
    w = np.array(wells, dtype=float)
    real = Realizations([K], [n], [b], w[:, 3].reshape(1, -1), np.array(coef, float))
    om = O.OracleModel(0.0, K, n, b, w, 0.0, 0.0, coef)
    return np.ascontiguousarray(w[:, :2]), real, om


@pytest.mark.parametrize("case", ["on_well", "step_cap", "zero_duration", "forward",
                                  "unconfined_dry", "unconfined_ok"])
def test_trace_status_cases_match_oracle(gpu_ctx, case):
    wells = [(0.0, 0.0, 0.25, 500.0), (30.0, 0.0, 0.25, 200.0), (-40.0, 25.0, 0.25, -150.0)]
    coef = (1e-5, -2e-5, 0.0, 1e-2, -5e-3, 2000.0)
    duration, confined, cap, start = 1000.0, True, 0, (1.25, 0.0)
    if case == "on_well":
        start = (30.0, 0.0)
    elif case == "step_cap":
        cap = 7
    elif case == "zero_duration":
        duration = 0.0
    elif case == "forward":
        duration = -500.0
    elif case == "unconfined_dry":
        coef, confined = (0, 0, 0, 0, 0, -1e9), False
    elif case == "unconfined_ok":
        confined = False
    wxy, real, om = _one_model(wells, coef=coef)
    verts, nverts, status = trace_vertices(wxy, 0.0, 0.0, real, np.array([start]), duration, 1.0,
                                           10.0, confined, cap)
    ov, res = om.backtrace(start[0], start[1], duration, 1.0, 10.0, confined=confined,
                           max_steps=cap)
    assert status[0] == res.status
    assert nverts[0] == res.nverts
    assert np.array_equal(verts[0, :nverts[0]], ov)


DIV_CASES = {
    # the start point shares its x (then its y) with a well: zero numerators
    "same_x_as_well": ([(0.0, 0.0, 0.25, 500.0), (30.0, 0.0, 0.25, 200.0),
                        (-40.0, 25.0, 0.25, -150.0)], (30.0, 7.5)),
    "same_y_as_well": ([(0.0, 0.0, 0.25, 500.0), (30.0, 0.0, 0.25, 200.0),
                        (-40.0, 25.0, 0.25, -150.0)], (-12.0, 25.0)),
    # coordinates beyond 2^60: every evaluation takes the scaled divisions
    "beyond_2p60": ([(1.5e18, 0.0, 0.25, 500.0), (1.5e18 + 4096.0, 2048.0, 0.25, 200.0)],
                    (1.5e18 + 1024.0, 512.0)),
    # a discharge below 2^-100 and a well coordinate below 2^-200: the
    # realization takes the scaled divisions
    "tiny_strength": ([(0.0, 0.0, 0.25, 500.0), (30.0, 0.0, 0.25, 1e-40)], (1.25, 0.5)),
    "tiny_coordinate": ([(0.0, 0.0, 0.25, 500.0), (1e-70, 40.0, 0.25, 200.0)], (1.25, 0.5)),
}


@pytest.mark.parametrize("case", sorted(DIV_CASES))
@pytest.mark.parametrize("confined", [True, False], ids=["confined", "unconfined"])
def test_fp64_division_paths_bit_exact_vs_oracle(gpu_ctx, case, confined):
    """The FP64 tracer divides with the short sequence where no operand can
    need div_scale's scaling and with the full one elsewhere; both give the
    oracle's vertices bit for bit, including zero numerators (a point on a
    well's grid line) and points, wells and discharges outside the short
    sequence's range."""
    wells, start = DIV_CASES[case]
    coef = (1e-5, -2e-5, 0.0, 1e-2, -5e-3, 2000.0)
    xo, yo = wells[0][0], wells[0][1]
    w = np.array(wells, dtype=float)
    real = Realizations([10.0], [0.25], [20.0], w[:, 3].reshape(1, -1), np.array(coef, float))
    om = O.OracleModel(0.0, 10.0, 0.25, 20.0, w, xo, yo, coef)
    verts, nverts, status = trace_vertices(np.ascontiguousarray(w[:, :2]), xo, yo, real,
                                           np.array([start]), 400.0, 1.0, 10.0, confined, 0)
    ov, res = om.backtrace(start[0], start[1], 400.0, 1.0, 10.0, confined=confined)
    assert status[0] == res.status
    assert nverts[0] == res.nverts > 3
    assert np.array_equal(verts[0, :nverts[0]], ov)
    assert np.array_equal(np.signbit(verts[0, :nverts[0]]), np.signbit(ov))


def test_empty_batches_are_noops(gpu_ctx):
    s, d = small_site(nreal=2, npaths=4)
    _, _, coef = oracle_fit_draw(s, d)
    pf = s.fixed_field()
    real = Realizations(d["K"][:0], d["n"][:0], d["b"][:0], d["q"][:0], coef[:0])
    res = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, start_points(s.xt, s.yt, s.rt, 4),
                               s.duration, s.tol, s.maxstep, s.umbra)
    assert res.counts.sum() == 0
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    res = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, np.zeros((0, 2)), s.duration,
                               s.tol, s.maxstep, s.umbra)
    assert res.counts.sum() == 0


BITMAP_CASES = [
    ("auto", _lib.BITMAP_AUTO, 0),           # 1200 x 1100 nodes: tiled (41250 flat words)
    ("tiled_overflow", _lib.BITMAP_TILED, 3),   # 3 LDS slots: most tiles live in HBM
    ("tiled_one_slot", _lib.BITMAP_TILED, 1),
    ("hbm_flat", _lib.BITMAP_HBM_FLAT, 0),
]


@pytest.mark.parametrize("name,mode,slots", BITMAP_CASES, ids=[c[0] for c in BITMAP_CASES])
@pytest.mark.parametrize("precision", [_lib.FP64_EXACT, _lib.FP32], ids=["fp64", "fp32"])
def test_large_grid_bitmap_placements(gpu_ctx, name, mode, slots, precision):
    """A 1200 x 1100 lattice (41250 bitmap words) exceeds the flat LDS budget.
    Tiled bitmaps (with or without overflow into HBM) and flat HBM bitmaps
    give the oracle's counts bit for bit (FP64) or the FP64 path's counts
    within the FP32 tolerance."""
    print(f"case {name} precision {precision}", flush=True)
    s, d = small_site(nreal=5, npaths=40)
    _, _, coef = oracle_fit_draw(s, d)
    pf = fixed_field(s.xt, s.yt, 5.0, 1100, 1200)
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    res = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol,
                               s.maxstep, 10.0, precision=precision, bitmap_mode=mode,
                               tile_slots=slots)
    assert res.passes == 1
    assert res.counts.sum() == res.stats.visited
    if precision == _lib.FP64_EXACT:
        s.umbra = 10.0
        ocounts, ost = oracle_chronicle(s, d, coef, pf)
        assert ost[6] == 0
        assert np.array_equal(res.counts, ocounts)
    else:
        ref = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                   s.tol, s.maxstep, 10.0, precision=precision,
                                   bitmap_mode=_lib.BITMAP_HBM_FLAT)
        assert np.array_equal(res.counts, ref.counts)


def test_small_grid_tiled_matches_flat(gpu_ctx):
    """The same realizations on the config-1 lattice (flat LDS bitmap by
    default) through the tiled bitmap, with and without overflow into HBM:
    identical counts."""
    s = sites.config1(6)
    d = sites.draw(s, seed=5)
    _, _, coef = oracle_fit_draw(s, d)
    pf = s.fixed_field()
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    out = {}
    for mode, slots in ((_lib.BITMAP_LDS_FLAT, 0), (_lib.BITMAP_TILED, 0), (_lib.BITMAP_TILED, 2)):
        print(f"mode {mode} slots {slots}", flush=True)
        out[(mode, slots)] = chronicle_into_field(
            pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol, s.maxstep, s.umbra,
            precision=_lib.FP32, bitmap_mode=mode, tile_slots=slots).counts
    base = out[(_lib.BITMAP_LDS_FLAT, 0)]
    assert base.sum() > 0
    for k, v in out.items():
        assert np.array_equal(v, base), k


def test_config1_subset_bit_exact_and_invariants(gpu_ctx):
    s = sites.config1(8)
    d = sites.draw(s, seed=2)
    _, _, coef = oracle_fit_draw(s, d)
    pf = s.fixed_field()
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    res = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol,
                               s.maxstep, s.umbra)
    ocounts, ost = oracle_chronicle(s, d, coef, pf)
    assert np.array_equal(res.counts, ocounts)
    assert (res.stats.attempted, res.stats.accepted, res.stats.visited) == (ost[0], ost[1], ost[7])
    # invariants: a node counts at most once per realization; the nodes within
    # umbra of the start ring are visited by every realization
    assert res.counts.max() <= s.nrealizations
    assert res.counts.sum() == res.stats.visited
    jt = int(round((s.xt - pf.xmin) / pf.deltax))
    it = int(round((s.yt - pf.ymin) / pf.deltay))
    assert res.counts[it, jt] == s.nrealizations


def test_config1_full_size_properties(gpu_ctx):
    """The headline workload at its full size (10^4 realizations x 1000
    particles, 400^2 nodes), through properties that hold at any size: FP32
    within 0.01 of FP64-exact at every node; node counts additive over
    realization subsets (two halves, both tracers); the pool schedule equal to
    the workgroup one; counts bounded by R, summing to the visited total and
    equal to R at the target well."""
    from onekapy_amd import fit_and_draw
    s = sites.config1()
    R = s.nrealizations
    d = sites.draw(s, seed=11)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"])
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)

    def run(sl, **kw):
        real = Realizations(d["K"][sl], d["n"][sl], d["b"][sl], d["q"][sl], coef[sl])
        pf = s.fixed_field()
        return chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                    s.tol, s.maxstep, s.umbra, **kw)

    halves = (slice(0, R // 2), slice(R // 2, R))
    out = {}
    for prec in (_lib.FP64_EXACT, _lib.FP32):
        full = run(slice(None), precision=prec)
        assert full.stats.escaped == 0
        assert list(full.stats.status_count) == [R * s.npaths, 0, 0, 0]
        parts = [run(h, precision=prec).counts for h in halves]
        assert np.array_equal(full.counts, parts[0] + parts[1])
        assert full.counts.max() <= R and full.counts.sum() == full.stats.visited
        pf = s.fixed_field()
        jt = int(round((s.xt - pf.xmin) / pf.deltax))
        it = int(round((s.yt - pf.ymin) / pf.deltay))
        assert full.counts[it, jt] == R
        out[prec] = full
    pool = run(slice(None), precision=_lib.FP32, schedule=_lib.SCHED_POOL)
    assert np.array_equal(pool.counts, out[_lib.FP32].counts)
    err = np.abs(out[_lib.FP32].counts - out[_lib.FP64_EXACT].counts) / R
    assert err.max() <= 0.01
    assert err.sum() <= 0.01 * (out[_lib.FP64_EXACT].counts > 0).sum()


def test_config4_shape_properties(gpu_ctx):
    """The large regional stress case at its real shape (200 wells, 10 %
    injection, 200 observations, 2000^2 nodes at 10 m, 256 particles, 30
    years) on 1,000 of its realizations: FP32 within 0.01 of FP64-exact at
    every node, with every well summed directly and with far-well expansions,
    the pool schedule equal to the workgroup one, counts additive over
    realization halves and bounded by R."""
    from onekapy_amd import fit_and_draw
    R = 1000
    s = sites.config4(R)
    d = sites.draw(s, seed=4)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"])
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)

    def run(sl, **kw):
        real = Realizations(d["K"][sl], d["n"][sl], d["b"][sl], d["q"][sl], coef[sl])
        pf = s.fixed_field()
        return chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                    s.tol, s.maxstep, s.umbra, **kw)

    print("fp32 workgroup", flush=True)
    wg = run(slice(None), precision=_lib.FP32, schedule=_lib.SCHED_WORKGROUP)
    print("fp32 pool", flush=True)
    pool = run(slice(None), precision=_lib.FP32, schedule=_lib.SCHED_POOL, far_order=-1)
    assert np.array_equal(pool.counts, wg.counts)
    assert list(pool.stats.status_count) == list(wg.stats.status_count)
    print("fp32 pool, far-well expansions: halves and whole", flush=True)
    far = run(slice(None), precision=_lib.FP32, schedule=_lib.SCHED_POOL)
    halves = [run(h, precision=_lib.FP32, schedule=_lib.SCHED_POOL).counts
              for h in (slice(0, R // 2), slice(R // 2, R))]
    assert np.array_equal(far.counts, halves[0] + halves[1])
    assert far.counts.max() <= R and far.counts.sum() == far.stats.visited
    assert wg.counts.max() <= R and wg.counts.sum() == wg.stats.visited
    print("fp64", flush=True)
    f64 = run(slice(None), precision=_lib.FP64_EXACT)
    for name, res in (("direct", wg), ("far", far)):
        err = np.abs(res.counts - f64.counts) / R
        print(f"{name}: max |p32 - p64| {err.max():.4f}, L1 {err.sum():.3f}", flush=True)
        assert err.max() <= 0.01
        assert err.sum() <= 0.01 * (f64.counts > 0).sum()


def test_window_marks_match_oracle_per_realization(gpu_ctx):
    from onekapy_amd.surrogate import cell_probabilities, trace_marks, _window
    s, d = small_site(nreal=5, npaths=48)
    _, _, coef = oracle_fit_draw(s, d)
    pf = s.fixed_field()
    win = _window(30, 35, 3, 12)          # 12 x 12 cells of 3 x 3 nodes
    marks, st = trace_marks(s, d, coef, win, _lib.FP64_EXACT)
    for r in range(5):
        dr = {k: v[r:r + 1] for k, v in d.items()}
        oc, _ = oracle_chronicle(s, dr, coef[r:r + 1], pf)
        want = np.rint(cell_probabilities(oc, 1, win) * 9).astype(np.uint8)
        assert np.array_equal(marks[r], want), r
    assert marks.max() <= 9


def test_fp32_error_config1_subset(gpu_ctx):
    s = sites.config1(200)
    d = sites.draw(s, seed=3)
    from onekapy_amd import fit_and_draw
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"])
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    p = {}
    for prec in (_lib.FP64_EXACT, _lib.FP32):
        pf = s.fixed_field()
        p[prec] = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                       s.tol, s.maxstep, s.umbra, precision=prec).counts / 200.0
    err = np.abs(p[_lib.FP32] - p[_lib.FP64_EXACT])
    assert err.max() <= 0.01
    assert err.sum() <= 0.01 * (p[_lib.FP64_EXACT] > 0).sum()


def _pool_vs_workgroup(s, d, coef, pf, umbra, **pool):
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    args = (pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol, s.maxstep, umbra)
    wg = chronicle_into_field(*args, precision=_lib.FP32, schedule=_lib.SCHED_WORKGROUP)
    pl = chronicle_into_field(*args, precision=_lib.FP32, schedule=_lib.SCHED_POOL, **pool)
    assert np.array_equal(pl.counts, wg.counts)
    assert np.array_equal(pl.trace_status, wg.trace_status)
    for k in ("attempted", "accepted", "visited", "escaped", "candidates"):
        assert getattr(pl.stats, k) == getattr(wg.stats, k), k
    assert list(pl.stats.status_count) == list(wg.stats.status_count)
    return pl


POOL_CASES = [
    ("default", {}),                                            # chunks of 64 particles
    ("whole_realizations", dict(pool_chunk=1000)),              # one chunk: direct flush
    ("odd_chunks_one_slot", dict(pool_chunk=7, pool_slots=1)),
    ("tiny_pool", dict(pool_slots=4, pool_waves_per_cu=24)),   # ~5 LDS tiles per slot: HBM
]


@pytest.mark.parametrize("name,pool", POOL_CASES, ids=[c[0] for c in POOL_CASES])
def test_pool_schedule_matches_workgroup_config1(gpu_ctx, name, pool):
    """The persistent wave-pool schedule (work stealing over live particles)
    traces the same particles, so counts, statuses and statistics equal the
    workgroup schedule's."""
    print(f"pool case {name}", flush=True)
    s = sites.config1(40)
    d = sites.draw(s, seed=7)
    _, _, coef = oracle_fit_draw(s, d)
    _pool_vs_workgroup(s, d, coef, s.fixed_field(), s.umbra, **pool)


def test_pool_schedule_rejects_what_does_not_fit(gpu_ctx):
    s = sites.config1(2)
    d = sites.draw(s, seed=7)
    _, _, coef = oracle_fit_draw(s, d)
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    args = (s.fixed_field(), s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol,
            s.maxstep, s.umbra)
    with pytest.raises(_lib.OkaError, match="do not fit in LDS"):
        chronicle_into_field(*args, precision=_lib.FP32, schedule=_lib.SCHED_POOL,
                             pool_slots=8, pool_waves_per_cu=32)
    with pytest.raises(_lib.OkaError, match="do not fit in LDS"):
        chronicle_into_field(*args, precision=_lib.FP64_EXACT, schedule=_lib.SCHED_POOL,
                             pool_slots=16, pool_waves_per_cu=32)


def test_pool_schedule_many_tiles_and_uneven_lifetimes(gpu_ctx):
    """A wide umbra on a fine 1200 x 1100 lattice marks more tiles than a
    slot's map holds (unmapped tiles in HBM), and mixed realization sizes
    (npaths 40) with a long duration give uneven lifetimes."""
    s, d = small_site(nreal=12, npaths=40)
    _, _, coef = oracle_fit_draw(s, d)
    pf = fixed_field(s.xt, s.yt, 5.0, 1100, 1200)
    for chunk in (40, 16, 5):
        print(f"chunk {chunk}", flush=True)
        pl = _pool_vs_workgroup(s, d, coef, pf, 300.0, pool_slots=3, pool_chunk=chunk)
        assert pl.counts.max() == 12


def _raster_polyline(grid, xy, umbra, flags):
    ctx = _lib.context()
    marks = np.zeros((grid.nrows, grid.ncols), dtype=np.uint8)
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    ctx.check(ctx.lib.oka_rasterize(ctx.h, grid, xy.shape[0], _lib.ptr(xy), float(umbra),
                                    _lib.ptr(marks), flags))
    return marks.astype(bool)


@pytest.mark.parametrize("spacing,umbra", [(20.0, 40.0), (10.0, 20.0), (40.0, 15.0)])
def test_fp32_rasterizer_matches_exact_but_at_the_umbra_circle(gpu_ctx, spacing, umbra):
    """The FP32 tracers' rasterizer (row chords of the umbra capsule) marks the
    same nodes as the FP64-exact one (per-node test, itself bit-exact against
    the oracle), except nodes whose exact distance to the polyline is within
    rounding of umbra.  The polyline mixes steps shorter than a cell, long
    jumps (> 2 umbra), exact axis-parallel and zero-length segments, and parts
    outside the grid."""
    rng = np.random.default_rng(11)
    nrows, ncols = 180, 170
    x0, y0 = -1234.5, 777.25
    grid = _lib.make_grid(x0, y0, x0 + (ncols - 1) * spacing, y0 + (nrows - 1) * spacing,
                          spacing, spacing, nrows, ncols)
    pts = [(x0 + 0.5 * ncols * spacing, y0 + 0.5 * nrows * spacing)]
    for k in range(6000):
        x, y = pts[-1]
        kind = k % 6
        if kind == 0:                      # long jump anywhere, partly off the grid
            x = x0 + rng.uniform(-0.1, 1.1) * ncols * spacing
            y = y0 + rng.uniform(-0.1, 1.1) * nrows * spacing
        elif kind == 1:                    # exactly horizontal or vertical
            if rng.random() < 0.5:
                x = x + rng.uniform(-3, 3) * spacing
            else:
                y = y + rng.uniform(-3, 3) * spacing
        elif kind == 2:                    # zero length
            pass
        else:                              # tracer-like steps up to 10 m
            th = rng.uniform(0, 2 * np.pi)
            ds = rng.uniform(0, 10.0) if kind < 5 else rng.uniform(0, 1e-3)
            x, y = x + ds * np.cos(th), y + ds * np.sin(th)
        pts.append((x, y))
    xy = np.array(pts)
    exact = _raster_polyline(grid, xy, umbra, 0)
    fp32 = _raster_polyline(grid, xy, umbra, _lib.RASTER_FP32)
    assert exact.sum() > 10000
    diff = np.argwhere(exact != fp32)
    assert len(diff) <= 1e-4 * exact.sum()
    # float coordinates relative to the grid centre resolve ~5e-4 m here
    tol = 64 * np.spacing(np.float32(0.6 * max(nrows, ncols) * spacing))
    for i, j in diff:
        cx, cy = x0 + j * spacing, y0 + i * spacing
        d = min(np.sqrt(O.distancesquared(ax, ay, bx, by, cx, cy))
                for (ax, ay), (bx, by) in zip(xy[:-1], xy[1:])
                if np.hypot(ax - bx, ay - by) > 0)
        assert abs(d - umbra) <= tol, (i, j, d)
