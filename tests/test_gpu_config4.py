"""The large regional stress case (200 wells, 10 % of them injecting, 200
observations, 2000^2 nodes at 10 m, 256 particles, 30-year backtraces) under
the oracle: the FP64-exact GPU path equals the oracle bit for bit on whole
realizations -- fit and MVN draw, node counts (tiled and flat HBM bitmaps),
attempted and accepted steps, visited nodes, and trace statuses including
step-capped traces -- and the FP32 pool tracer agrees with it trace by trace."""
import numpy as np
import pytest

from onekapy_amd import Realizations, _lib, chronicle_into_field, fit_and_draw, sites, start_points
from helpers import oracle_chronicle, oracle_fit_draw

pytestmark = pytest.mark.gpu

# realizations 0..11 of the seed-4 draw of 96: with a 5,000-step cap, 3 and 5
# hold 154 and 77 capped traces (traces spiralling into injection wells)
NREAL_DRAW, SEED, FIRST = 96, 4, 12
CAP = 5000


def _subset():
    s = sites.config4(NREAL_DRAW)
    d = sites.draw(s, seed=SEED)
    d = {k: v[:FIRST] for k, v in d.items()}
    s.nrealizations = FIRST
    return s, d


def _run(s, d, coef, **kw):
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    return chronicle_into_field(s.fixed_field(), s.wells_xy(), s.xt, s.yt, real, starts,
                                s.duration, s.tol, s.maxstep, s.umbra, max_steps=CAP, **kw)


@pytest.fixture(scope="module")
def cfg4_oracle():
    s, d = _subset()
    _, _, coef = oracle_fit_draw(s, d)
    import oracle as O
    pf = s.fixed_field()
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    opts = O.trace_opts(s.duration, s.tol, s.maxstep, True, O.MATH_CR, CAP)
    g = O.make_grid(pf.xmin, pf.ymin, pf.xmax, pf.ymax, pf.deltax, pf.deltay, pf.nrows, pf.ncols)
    counts, st = O.chronicle_fixed(s.wells_xy(), s.xt, s.yt, s.base, d["K"], d["n"], d["b"],
                                   d["q"], coef, starts, s.umbra, opts, g)
    return s, d, coef, counts, st


def test_config4_fit_and_draw_bit_exact(gpu_ctx, cfg4_oracle):
    s, d, coef, _, _ = cfg4_oracle
    _, _, gcoef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                               d["q"], d["z"])
    assert np.array_equal(gcoef, coef)


CFG4_RUNS = [("tiled", dict(bitmap_mode=_lib.BITMAP_AUTO)),
             ("hbm_flat", dict(bitmap_mode=_lib.BITMAP_HBM_FLAT)),
             ("pool", dict(schedule=_lib.SCHED_POOL))]


@pytest.mark.parametrize("name,kw", CFG4_RUNS, ids=[c[0] for c in CFG4_RUNS])
def test_config4_fp64_exact_matches_oracle(gpu_ctx, cfg4_oracle, name, kw):
    s, d, coef, ocounts, ost = cfg4_oracle
    print(f"cfg4 fp64-exact, {name}", flush=True)
    res = _run(s, d, coef, precision=_lib.FP64_EXACT, **kw)
    assert res.passes == 1 and ost[6] == 0
    assert np.array_equal(res.counts, ocounts)
    st = res.stats
    assert (st.attempted, st.accepted, st.visited) == (ost[0], ost[1], ost[7])
    assert list(st.status_count) == list(ost[2:6])
    assert st.status_count[_lib.TRACE_STEP_CAP] > 100          # traces that hit the cap
    assert np.array_equal(np.bincount(res.trace_status, minlength=4), ost[2:6])


def test_config4_fp32_pool_agrees_with_fp64_exact(gpu_ctx, cfg4_oracle):
    """On the same 12 realizations the FP32 pool tracer ends the same traces the
    same way (to 1 % of the traces: a trace near the cap may need a few more or
    fewer steps in FP32) and stays within the capture-probability error bound
    on average over the visited nodes."""
    s, d, coef, ocounts, ost = cfg4_oracle
    f64 = _run(s, d, coef, precision=_lib.FP64_EXACT)
    pool = _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_POOL, far_order=-1)
    wg = _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_WORKGROUP)
    assert np.array_equal(pool.counts, wg.counts)
    assert np.array_equal(pool.trace_status, wg.trace_status)
    mism = int((pool.trace_status != f64.trace_status).sum())
    print(f"per-trace status mismatches FP32 vs FP64: {mism} of {len(f64.trace_status)}",
          flush=True)
    assert mism <= 0.01 * len(f64.trace_status)
    R = len(d["K"])
    err = np.abs(pool.counts - f64.counts) / R
    print(f"max |p32 - p64| {err.max():.4f}, L1 {err.sum():.4f}, visited nodes "
          f"{(f64.counts > 0).sum()}", flush=True)
    assert err.sum() <= 0.01 * (f64.counts > 0).sum()


@pytest.mark.parametrize("order", [0, 8, 16], ids=["auto", "order8", "order16"])
def test_config4_far_expansions_agree_with_fp64_exact(gpu_ctx, cfg4_oracle, order):
    """The FP32 pool tracer with the wells beyond the 3 x 3 boxes around each
    point summed as local expansions (the automatic choice, order 16, and
    order 8) ends the same traces the same way as FP64-exact to 1 % of the traces
    and keeps the capture-probability error within the bound on average over
    the visited nodes; its traces stay close to the direct FP32 sum's."""
    s, d, coef, ocounts, ost = cfg4_oracle
    f64 = _run(s, d, coef, precision=_lib.FP64_EXACT)
    direct = _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_POOL, far_order=-1)
    far = _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_POOL, far_order=order)
    assert not np.array_equal(far.counts, direct.counts)   # the expansions were used
    mism = int((far.trace_status != f64.trace_status).sum())
    R = len(d["K"])
    err = np.abs(far.counts - f64.counts) / R
    rel = abs(far.stats.attempted - f64.stats.attempted) / f64.stats.attempted
    print(f"order {order}: status mismatches {mism}, max |p - p64| {err.max():.4f}, "
          f"L1 {err.sum():.4f}, visited {(f64.counts > 0).sum()}, steps {rel:.2e}", flush=True)
    assert mism <= 0.01 * len(f64.trace_status)
    assert err.sum() <= 0.01 * (f64.counts > 0).sum()
    assert rel < 0.01
    # bounding boxes of the traces that end normally in both
    ok = (far.trace_status == 0) & (direct.trace_status == 0)
    dev = np.abs(far.bbox[ok] - direct.bbox[ok]).max(axis=1)
    print(f"bbox deviation from the direct sum: median {np.median(dev):.3g} m, "
          f"95th percentile {np.percentile(dev, 95):.3g} m", flush=True)
    assert np.median(dev) < 1.0


def test_config4_auto_schedule(gpu_ctx, cfg4_oracle):
    """OKA_SCHED_AUTO on the 200-well site: FP32 takes the pool with the
    far-well expansions (the same counts, statuses and steps as asking for
    them); with far_order = -1, or a forced bitmap placement, and in
    FP64-exact it takes the workgroup schedule (the same counts as asking for
    it, and FP64-exact equal to the oracle)."""
    s, d, coef, ocounts, ost = cfg4_oracle
    auto32 = _run(s, d, coef, precision=_lib.FP32)
    far = _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_POOL)
    assert np.array_equal(auto32.counts, far.counts)
    assert np.array_equal(auto32.trace_status, far.trace_status)
    assert auto32.stats.attempted == far.stats.attempted
    wg32 = _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_WORKGROUP)
    assert not np.array_equal(auto32.counts, wg32.counts)
    direct = _run(s, d, coef, precision=_lib.FP32, far_order=-1)
    assert np.array_equal(direct.counts, wg32.counts)
    tiled = _run(s, d, coef, precision=_lib.FP32, bitmap_mode=_lib.BITMAP_TILED)
    assert np.array_equal(tiled.counts, wg32.counts)
    auto64 = _run(s, d, coef, precision=_lib.FP64_EXACT)
    assert np.array_equal(auto64.counts, ocounts)


def test_config4_far_order_above_limit_is_rejected(gpu_ctx, cfg4_oracle):
    s, d, coef, _, _ = cfg4_oracle
    with pytest.raises(_lib.OkaError, match="far_order"):
        _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_POOL, far_order=17)


def test_config4_shared_far_tables_match_per_chunk_tables(gpu_ctx, cfg4_oracle):
    """The chunks of a realization share one far-well table, computed by the
    first of them into the realization's record (chunks of 64 and 32
    particles: 4 and 8 chunks per realization); a realization of one chunk
    (256 particles) computes its own.  The tables hold the same coefficients,
    so every trace ends the same: equal counts, statuses, steps and bounding
    boxes."""
    s, d, coef, _, _ = cfg4_oracle
    runs = {c: _run(s, d, coef, precision=_lib.FP32, schedule=_lib.SCHED_POOL, pool_chunk=c)
            for c in (256, 64, 32)}
    one = runs[256]
    for c in (64, 32):
        r = runs[c]
        print(f"chunk {c}: attempted {r.stats.attempted} against {one.stats.attempted}",
              flush=True)
        assert np.array_equal(r.counts, one.counts)
        assert np.array_equal(r.trace_status, one.trace_status)
        assert np.array_equal(r.bbox, one.bbox)
        assert (r.stats.attempted, r.stats.accepted, r.stats.visited) == (
            one.stats.attempted, one.stats.accepted, one.stats.visited)


def test_config4_shared_far_tables_across_record_reuse(gpu_ctx):
    """Shared far-well tables when records are reused: 8,000 realizations of
    64 particles in chunks of 16 (4 chunks, one shared table each) on 1 wave
    per compute unit, whose 2,048 records each serve about four realizations in
    turn, against chunks of 64 (one chunk, a table per slot).  A table read
    before its realization's coefficients reached the reader would change
    traces; every count, status, step and bounding box is equal."""
    s = sites.config4(8000)
    d = sites.draw(s, seed=5)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"], moments=False)
    real = Realizations(d["K"], d["n"], d["b"], d["q"], coef)
    starts = start_points(s.xt, s.yt, s.rt, 64)
    out = {}
    for chunk in (16, 64):
        out[chunk] = chronicle_into_field(s.fixed_field(), s.wells_xy(), s.xt, s.yt, real, starts,
                                          s.duration, s.tol, s.maxstep, s.umbra,
                                          precision=_lib.FP32, max_steps=CAP,
                                          schedule=_lib.SCHED_POOL, pool_slots=2,
                                          pool_waves_per_cu=1, pool_chunk=chunk)
        print(f"chunk {chunk}: attempted {out[chunk].stats.attempted}", flush=True)
    a, b = out[16], out[64]
    assert np.array_equal(a.counts, b.counts)
    assert np.array_equal(a.trace_status, b.trace_status)
    assert np.array_equal(a.bbox, b.bbox)
    assert (a.stats.attempted, a.stats.accepted, a.stats.visited) == (
        b.stats.attempted, b.stats.accepted, b.stats.visited)
