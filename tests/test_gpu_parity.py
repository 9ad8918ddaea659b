"""GPU parity: liboka (HIP, gfx950) against the oracle on identical inputs.

Bar: bit-exact for the FP64-exact path (fit, draw, vertices, counts, step
counts); the FP32 tracer is held to max |p32 - p64| <= 0.01.
"""
import math

import numpy as np
import pytest

import oracle as O
from onekapy_amd import (Model, ProbabilityField, Realizations, chronicle_into_field,
                         compute_backtrace, compute_capturezone, create_stochastic_capturezone,
                         fit_and_draw, sites, start_points, trace_vertices, _lib)
from helpers import golden, oracle_chronicle, oracle_fit_draw, oracle_grid, small_site

pytestmark = pytest.mark.gpu


def _kat_model():
    k = golden("reference_kats.json")["model"]
    return Model(k["base"], k["conductivity"], k["porosity"], k["thickness"],
                 [tuple(w) for w in k["wells"]], k["xo"], k["yo"], np.array(k["coef"]))


def test_model_kats_on_gpu(gpu_ctx):
    g = golden("reference_kats.json")
    m = _kat_model()
    assert np.isclose(m.compute_potential(*g["potential"]["xy"]), g["potential"]["value"], rtol=1e-6)
    assert np.isclose(m.compute_head(*g["head"]["xy"]), g["head"]["value"], rtol=1e-6)
    assert np.allclose(m.compute_discharge(*g["discharge"]["xy"]), g["discharge"]["value"],
                       rtol=1e-6)
    for case in g["velocity"]["cases"]:
        assert np.allclose(m.compute_velocity(*case["xy"]), case["value"], rtol=1e-6)


def test_model_eval_bit_exact_vs_oracle(gpu_ctx):
    m = _kat_model()
    om = O.OracleModel(m.base, m.conductivity, m.porosity, m.thickness, m.wells, m.xo, m.yo,
                       m.coef)
    rng = np.random.default_rng(3)
    x = rng.uniform(-500, 800, 300)
    y = rng.uniform(-500, 800, 300)
    qx, qy = m.compute_discharge(x, y)
    vx, vy = m.compute_velocity_confined(x, y)
    pot = m.compute_potential(x, y)
    for i in range(len(x)):
        assert (qx[i], qy[i]) == tuple(om.discharge(x[i], y[i]))
        assert (vx[i], vy[i]) == tuple(om.velocity_confined(x[i], y[i]))
        assert pot[i] == om.potential(x[i], y[i], O.MATH_CR)


def _same_bits(a, b):
    return np.float64(a).tobytes() == np.float64(b).tobytes() or (math.isnan(a) and math.isnan(b))


def test_model_eval_edge_points_bit_exact_vs_oracle(gpu_ctx):
    """The FP64 discharge divides each well's x and y terms by the same r^2
    with one shared reciprocal refinement when the operands allow it, and by
    two plain divisions otherwise.  Both paths must give the oracle's bits.
    One well per model and no regional gradient, so each result is that
    well's quotient: zero numerators (points level with the well, q = 0),
    extreme exponents (q = 1e-300 and 1e300, points 1e-150 m from the well or
    1e7 m away, denormal coordinates), every sign, and the confined
    velocity's second shared division by b n."""
    rng = np.random.default_rng(8)
    base_pts = [(0.0, 7.0), (13.0, 0.0), (1e-150, 0.0), (0.0, 1e-150), (1e-150, 2e-150),
                (1e7, -1e7), (-3.5, 1e-9), (5e-324, 250.0), (1e-300, -1e-300), (2.0, 2.0)]
    nchecked = 0
    for q in (500.0, 0.0, -730.5, 1e-300, 1e300, -2.5e-20, 6e150):
        for xw, yw in ((0.0, 0.0), (100.0, -250.0)):
            m = Model(0.0, 10.0, 0.25, 20.0, [(xw, yw, 0.25, q)], 0.0, 0.0,
                      np.array([0.0, 0.0, 0.0, 0.0, 0.0, 2000.0]))
            om = O.OracleModel(m.base, m.conductivity, m.porosity, m.thickness, m.wells,
                               m.xo, m.yo, m.coef)
            pts = [(xw + a, yw + b) for a, b in base_pts]
            pts += [(xw + a, yw + b) for a, b in rng.uniform(-400, 400, (20, 2))]
            x = np.array([p[0] for p in pts])
            y = np.array([p[1] for p in pts])
            qx, qy = m.compute_discharge(x, y)
            vx, vy = m.compute_velocity_confined(x, y)
            for i in range(len(x)):
                ox, oy = om.discharge(x[i], y[i])
                assert _same_bits(qx[i], ox) and _same_bits(qy[i], oy), (q, xw, x[i], y[i])
                ox, oy = om.velocity_confined(x[i], y[i])
                assert _same_bits(vx[i], ox) and _same_bits(vy[i], oy), (q, xw, x[i], y[i])
                nchecked += 1
    assert nchecked == 7 * 2 * 30


def test_fit_kat_and_library_on_gpu(gpu_ctx):
    g = golden("reference_kats.json")["fit"]
    m = _kat_model()
    ev, cov = m.fit_regional_flow([tuple(o) for o in g["obs"]], *g["origin"])
    assert np.allclose(ev.ravel(), g["ev"], rtol=g["rtol"])
    assert np.allclose(cov, g["cov"], rtol=g["rtol"])
    # the reference's own library route: lstsq + inv on the weighted rows
    WA, Wb = m.construct_fit([tuple(o) for o in g["obs"]], *g["origin"])
    ev_np = np.linalg.lstsq(WA, Wb, rcond=-1)[0].ravel()
    cov_np = np.linalg.inv(WA.T @ WA)
    assert np.allclose(ev.ravel(), ev_np, rtol=1e-9, atol=0)
    assert np.allclose(cov, cov_np, rtol=1e-8, atol=0)


def test_fit_draw_bit_exact_vs_oracle(gpu_ctx):
    s, d = small_site(nreal=64)
    ev, cov, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                                 d["q"], d["z"])
    oev, ocov, ocoef = oracle_fit_draw(s, d)
    assert np.array_equal(ev, oev)
    assert np.array_equal(cov, ocov)
    assert np.array_equal(coef, ocoef)


def _site_real(s, d, coef):
    return Realizations(d["K"], d["n"], d["b"], d["q"], coef)


def test_vertices_bit_exact_vs_oracle(gpu_ctx):
    s, d = small_site(nreal=4, npaths=24)
    _, _, coef = oracle_fit_draw(s, d)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    verts, nverts, status = trace_vertices(s.wells_xy(), s.xt, s.yt, _site_real(s, d, coef),
                                           starts, s.duration, s.tol, s.maxstep)
    wxy = s.wells_xy()
    for r in range(4):
        om = O.OracleModel(s.base, d["K"][r], d["n"][r], d["b"][r],
                           np.c_[wxy, np.full(len(wxy), 0.25), d["q"][r]], s.xt, s.yt, coef[r])
        for p in range(s.npaths):
            ov, res = om.backtrace(starts[p, 0], starts[p, 1], s.duration, s.tol, s.maxstep)
            k = r * s.npaths + p
            assert nverts[k] == res.nverts
            assert status[k] == res.status == O.TRACE_OK
            assert np.array_equal(verts[k, :nverts[k]], ov)


def test_chronicle_fixed_grid_cfg0_bit_exact(gpu_ctx):
    s = sites.config0()
    d = sites.draw(s, seed=11)
    _, _, coef = oracle_fit_draw(s, d)
    pf = s.fixed_field()
    real = _site_real(s, d, coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    res = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol,
                               s.maxstep, s.umbra)
    ocounts, ost = oracle_chronicle(s, d, coef, pf)
    assert res.passes == 1 and ost[6] == 0
    assert np.array_equal(res.counts, ocounts)
    assert res.stats.attempted == ost[0]
    assert res.stats.accepted == ost[1]
    assert res.stats.visited == ost[7]
    assert ocounts.max() > 0


def test_stochastic_history_matches_sequential_oracle(gpu_ctx):
    """create_stochastic_capturezone from a fresh 3x3 field (two-pass history)
    equals the reference's sequential rasterize/register loop."""
    s, d = small_site(nreal=6, npaths=20)
    pf = create_stochastic_capturezone(0, s.npaths, s.duration, s.nrealizations, s.base,
                                       s.c_dist, s.p_dist, s.t_dist, s.wells, s.observations,
                                       s.spacing, s.umbra, True, s.tol, s.maxstep, seed=5)
    from onekapy_amd.stochastic import draw_realizations
    d = draw_realizations(s.wells, s.c_dist, s.p_dist, s.t_dist, s.nrealizations,
                          np.random.default_rng(5))
    _, _, coef = oracle_fit_draw(s, d)
    of = O.OracleField(s.spacing, s.spacing, s.xt, s.yt)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    opts = O.trace_opts(s.duration, s.tol, s.maxstep, True, O.MATH_CR)
    wxy = s.wells_xy()
    for r in range(s.nrealizations):
        om = O.OracleModel(s.base, d["K"][r], d["n"][r], d["b"][r],
                           np.c_[wxy, np.full(len(wxy), 0.25), d["q"][r]], s.xt, s.yt, coef[r])
        of.capturezone(om, starts, s.umbra, 1.0, opts)
    assert (pf.nrows, pf.ncols) == of.shape
    assert (pf.xmin, pf.xmax, pf.ymin, pf.ymax) == (of.s.xmin, of.s.xmax, of.s.ymin, of.s.ymax)
    assert np.array_equal(pf.pgrid, of.pgrid)
    assert pf.total_weight == of.s.total_weight == float(s.nrealizations)


def test_compute_capturezone_dropin(gpu_ctx):
    """compute_capturezone into an existing field, twice, equals the oracle's
    sequential reference loop, including growth between calls."""
    s, d = small_site(nreal=2, npaths=16)
    _, _, coef = oracle_fit_draw(s, d)
    pf = ProbabilityField(s.spacing, s.spacing, s.xt, s.yt)
    of = O.OracleField(s.spacing, s.spacing, s.xt, s.yt)
    opts = O.trace_opts(s.duration, s.tol, s.maxstep, True, O.MATH_CR)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    wxy = s.wells_xy()
    for r, w in zip(range(2), (1.0, 0.5)):
        wells = [(wxy[i, 0], wxy[i, 1], 0.25, d["q"][r][i]) for i in range(len(wxy))]
        m = Model(s.base, d["K"][r], d["n"][r], d["b"][r], wells, s.xt, s.yt, coef[r])
        compute_capturezone(s.xt, s.yt, s.rt, s.npaths, s.duration, pf, s.umbra, w, s.tol,
                            s.maxstep, m.velocity_function(True))
        om = O.OracleModel(s.base, d["K"][r], d["n"][r], d["b"][r], wells, s.xt, s.yt, coef[r])
        of.capturezone(om, starts, s.umbra, w, opts)
        assert (pf.nrows, pf.ncols) == of.shape
        assert np.array_equal(pf.pgrid, of.pgrid)


def test_compute_backtrace_dropin(gpu_ctx):
    m = _kat_model()
    m.coef = np.array([1e-4, 2e-4, 0.0, 1e-2, -1e-2, 500.0])
    fe = m.velocity_function(True)
    verts = compute_backtrace(150.0, 150.0, 100.0, 1.0, 10.0, fe)
    om = O.OracleModel(m.base, m.conductivity, m.porosity, m.thickness, m.wells, m.xo, m.yo,
                       m.coef)
    ov, res = om.backtrace(150.0, 150.0, 100.0, 1.0, 10.0)
    assert len(verts) == res.nverts
    assert np.array_equal(np.array(verts), ov)


def test_fp32_error_cfg0(gpu_ctx):
    s = sites.config0()
    d = sites.draw(s, seed=11)
    _, _, coef = oracle_fit_draw(s, d)
    real = _site_real(s, d, coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    out = {}
    for prec in (_lib.FP64_EXACT, _lib.FP32):
        pf = s.fixed_field()
        out[prec] = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                         s.tol, s.maxstep, s.umbra, precision=prec).counts
    p64 = out[_lib.FP64_EXACT] / s.nrealizations
    p32 = out[_lib.FP32] / s.nrealizations
    assert np.abs(p32 - p64).max() <= 0.01


def _dry_at_start(status, P):
    """Realizations all of whose traces ended truncated (AquiferError)."""
    return set(np.nonzero((status.reshape(-1, P) == _lib.TRACE_TRUNCATED).all(axis=1))[0])


UNCONF_CASES = [(0.0, _lib.BITMAP_AUTO, 0), (0.7, _lib.BITMAP_AUTO, 0), (1e6, _lib.BITMAP_AUTO, 0),
                (0.7, _lib.BITMAP_TILED, 1), (0.7, _lib.BITMAP_TILED, 3),
                (0.7, _lib.BITMAP_HBM_FLAT, 0)]


@pytest.mark.parametrize("shift,mode,slots", UNCONF_CASES,
                         ids=[f"{c[0]:g}-mode{c[1]}-slots{c[2]}" for c in UNCONF_CASES])
def test_fp32_unconfined_error_cfg0(gpu_ctx, shift, mode, slots):
    """The unconfined velocity (model.py:318-389) on cfg 0, whose potentials
    straddle K b^2 / 2 (the head is below the aquifer top in some realizations
    and above it in others).  F is lowered by shift * K b^2 / 2: at 0.7 a fifth
    of the realizations are dry at their start points (AquiferError, status
    truncated), at 10^6 every one is.  The FP64-exact mode is bit-exact against
    the oracle; the FP32 mode is within 0.01 of it in capture probability, ends
    the same traces the same way (per trace, to 0.1 %) and finds the same
    realizations dry at their start points.  The bitmap placements (flat LDS,
    tiled with 1 or 3 LDS slots, flat HBM) give the same FP32 counts."""
    print(f"unconfined shift {shift} bitmap {mode} slots {slots}", flush=True)
    s = sites.config0()
    d = sites.draw(s, seed=11)
    _, _, coef = oracle_fit_draw(s, d)
    coef[:, 5] -= shift * 0.5 * d["K"] * d["b"] ** 2
    real = _site_real(s, d, coef)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    P, ntr = len(starts), len(starts) * s.nrealizations
    g = oracle_grid(s.fixed_field())
    out, status = {}, {}
    for prec in (_lib.FP64_EXACT, _lib.FP32):
        pf = s.fixed_field()
        res = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                   s.tol, s.maxstep, s.umbra, confined=False, precision=prec,
                                   bitmap_mode=mode, tile_slots=slots)
        assert res.counts.shape == (g.nrows, g.ncols)   # no trace left the fixed grid
        out[prec], status[prec] = res.counts, res.trace_status
    if mode != _lib.BITMAP_AUTO:
        pf = s.fixed_field()
        flat = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                    s.tol, s.maxstep, s.umbra, confined=False,
                                    precision=_lib.FP32, bitmap_mode=_lib.BITMAP_LDS_FLAT)
        assert np.array_equal(out[_lib.FP32], flat.counts)
        assert np.array_equal(status[_lib.FP32], flat.trace_status)
    oc, ost = oracle_chronicle(s, d, coef, s.fixed_field(), confined=False)
    assert np.array_equal(out[_lib.FP64_EXACT], oc)
    n64 = np.bincount(status[_lib.FP64_EXACT], minlength=4)
    assert np.array_equal(n64, ost[2:6])
    n32 = np.bincount(status[_lib.FP32], minlength=4)
    assert (status[_lib.FP32] != status[_lib.FP64_EXACT]).sum() <= 0.001 * ntr
    assert _dry_at_start(status[_lib.FP32], P) == _dry_at_start(status[_lib.FP64_EXACT], P)
    p64 = out[_lib.FP64_EXACT] / s.nrealizations
    p32 = out[_lib.FP32] / s.nrealizations
    assert np.abs(p32 - p64).max() <= 0.01
    if shift == 0.0:
        assert not _dry_at_start(status[_lib.FP64_EXACT], P)
        # the unconfined field is not the confined one on this site
        pf = s.fixed_field()
        conf = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration,
                                    s.tol, s.maxstep, s.umbra, precision=_lib.FP32).counts
        assert np.abs(conf / s.nrealizations - p32).max() > 0.05
    elif shift == 0.7:
        assert 0 < n64[_lib.TRACE_TRUNCATED] < ntr
        assert 0 < len(_dry_at_start(status[_lib.FP64_EXACT], P)) < s.nrealizations
    else:
        assert n32[_lib.TRACE_TRUNCATED] == n64[_lib.TRACE_TRUNCATED] == ntr
        assert out[_lib.FP32].sum() == 0 and oc.sum() == 0


@pytest.mark.parametrize("precision", [_lib.FP64_EXACT, _lib.FP32], ids=["fp64", "fp32"])
def test_unconfined_window_marks_match_counts(gpu_ctx, precision):
    """Window marks of the unconfined field (F lowered as above, a fifth of the
    realizations dry): per realization they equal the cell sums of that
    realization's own node counts, and summed over realizations the counts of
    the chronicle launch."""
    from onekapy_amd.surrogate import trace_marks, _window
    s = sites.config0()
    s.nrealizations = 24
    s.confined = False
    d = sites.draw(s, seed=11)
    _, _, coef = oracle_fit_draw(s, d)
    coef[:, 5] -= 0.7 * 0.5 * d["K"] * d["b"] ** 2
    win = _window(20, 22, 3, 20)          # 20 x 20 cells of 3 x 3 nodes
    marks, st = trace_marks(s, d, coef, win, precision)
    starts = start_points(s.xt, s.yt, s.rt, s.npaths)
    total = np.zeros(win.nrows * win.ncols, np.int64)
    for r in range(s.nrealizations):
        pf = s.fixed_field()
        real = Realizations(d["K"][r:r + 1], d["n"][r:r + 1], d["b"][r:r + 1], d["q"][r:r + 1],
                            coef[r:r + 1])
        c = chronicle_into_field(pf, s.wells_xy(), s.xt, s.yt, real, starts, s.duration, s.tol,
                                 s.maxstep, s.umbra, confined=False, precision=precision).counts
        want = np.zeros((win.nrows * 3, win.ncols * 3), np.int64)
        want[:, :] = c[20:20 + win.nrows * 3, 22:22 + win.ncols * 3]
        want = want.reshape(win.nrows, 3, win.ncols, 3).sum(axis=(1, 3)).ravel()
        assert np.array_equal(marks[r].astype(np.int64), want), r
        total += want
    assert total.sum() > 0 and marks.sum() == total.sum()
