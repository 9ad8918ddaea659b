"""The learned control variate at node resolution: the tracer's per-cell node
masks against the oracle, oka_cv_node_sums against its parts (masks, network
probabilities and host sums), and the node estimate (a ProbabilityField over
the site's lattice) against a large plain run."""
import numpy as np
import pytest

from onekapy_amd import _lib, sites
from onekapy_amd.stochastic import draw_realizations, fit_and_draw
from onekapy_amd.surrogate import (PAIRS, Standardizer, SurrogateMLP, _node_image, _window,
                                   choose_window, control_variate_nodes, features, neighbours,
                                   node_reduce, node_sums, node_trace, trace_counts, trace_marks)
from helpers import oracle_chronicle, oracle_fit_draw, small_site

pytestmark = pytest.mark.gpu


def _mask_image(masks, win):
    """uint16 masks [R, cells] -> visited window nodes [R, rows, cols] (bool)."""
    nk = win.cell * win.cell
    bits = (masks[:, :, None] >> np.arange(nk, dtype=np.uint16)) & 1
    return _node_image(bits, win).astype(bool)


@pytest.mark.parametrize("precision", [_lib.FP64_EXACT, _lib.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("cell", [1, 3, 4])
def test_node_masks_are_the_visited_nodes(gpu_ctx, precision, cell):
    """Each realization's node masks are exactly its visited window nodes: the
    oracle's per-realization counts (FP64-exact), or the tracer's own count
    grid of that realization alone (FP32); their popcounts are the marks."""
    s, d = small_site(nreal=5, npaths=48)
    _, _, coef = oracle_fit_draw(s, d)
    pf = s.fixed_field()
    win = _window(30, 35, cell, 12)
    marks, masks, _ = trace_marks(s, d, coef, win, precision, masks=True)
    bits = (masks[:, :, None] >> np.arange(16, dtype=np.uint16)) & 1
    assert np.array_equal(bits.sum(2), marks)
    img = _mask_image(masks, win)
    h = w = 12 * cell
    for r in range(5):
        dr = {k: v[r:r + 1] for k, v in d.items()}
        if precision == _lib.FP64_EXACT:
            cnt, _ = oracle_chronicle(s, dr, coef[r:r + 1], pf)
        else:
            cnt, _ = trace_counts(s, dr, coef[r:r + 1], precision)
        want = np.zeros((h, w), bool)
        sub = cnt[30:30 + h, 35:35 + w] > 0
        want[:sub.shape[0], :sub.shape[1]] = sub
        assert np.array_equal(img[r], want), r
    assert img.any()
    with pytest.raises(_lib.OkaError, match="at most 4 x 4"):
        trace_marks(s, d, coef, _window(30, 35, 5, 12), precision, masks=True)


def test_node_sums_equal_their_parts(gpu_ctx):
    """oka_cv_node_sums against the same quantities built from trace_marks'
    node masks, the network's probabilities and sums on the host (cells and
    their 5 x 5 pair products, nodes and their 3 x 3 neighbouring cells), in
    three parts of ragged size across network batches; counts equal the
    plain tracer's."""
    s = sites.config0()
    rng = np.random.default_rng(5)
    R = 1300
    d = draw_realizations(s.wells, s.c_dist, s.p_dist, s.t_dist, R, rng)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"], moments=False)
    pc, _ = trace_counts(s, d, coef, _lib.FP32)
    win = choose_window(pc, cells=16, margin=1.1)
    win.cell = min(win.cell, 4)
    X = Standardizer(features(d, coef))(features(d, coef))
    net = SurrogateMLP(X.shape[1], 128, 256, max_batch=256, seed=4)
    counts = np.zeros_like(pc)
    sums, st = node_sums(net, s, d, coef, X, win, parts=3, counts=counts)
    assert np.array_equal(counts, pc)
    _, masks, _ = trace_marks(s, d, coef, win, _lib.FP32, masks=True)
    prob = net.predict(X).astype(np.float64)
    nk = win.cell * win.cell
    y = ((masks[:, :, None] >> np.arange(nk, dtype=np.uint16)) & 1).astype(np.float64)
    nb = neighbours(win)
    ncell = win.nrows * win.ncols
    pair = np.full((ncell, len(PAIRS)), -1)
    for a in range(ncell):
        r, c = divmod(a, win.ncols)
        for h, (dr, dc) in enumerate(PAIRS):
            if r + dr < win.nrows and 0 <= c + dc < win.ncols:
                pair[a, h] = (r + dr) * win.ncols + c + dc
    pz = np.concatenate([prob, np.zeros((R, 1))], axis=1)      # index -1: an absent cell
    b = [R * p // 3 for p in range(4)]
    for p in range(3):
        sl = slice(b[p], b[p + 1])
        np.testing.assert_allclose(sums["s"][p], prob[sl].sum(0), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(sums["ss"][p], (prob[sl][:, :, None] * pz[sl][:, pair]).sum(0),
                                   rtol=1e-12, atol=1e-12)
        assert np.array_equal(sums["y"][p], y[sl].sum(0))
        want = np.einsum("rak,ram->akm", y[sl], pz[sl][:, nb])
        np.testing.assert_allclose(sums["ys"][p], want, rtol=1e-12, atol=1e-12)
    assert sums["y"].sum() > 1000
    assert st.attempted > 0
    # the same sums in two calls: the set traced and kept on the device, then
    # reduced by the network on a context of its own
    counts2 = np.zeros_like(pc)
    ns, st2 = node_trace(s, d, coef, win, counts=counts2)
    net2 = SurrogateMLP(X.shape[1], 128, 256, max_batch=256, seed=4,
                        ctx=_lib.Context(_lib.context().device))
    sums2 = node_reduce(net2, ns, X, parts=3)
    ns.close()
    assert np.array_equal(counts2, pc)
    assert st2.attempted == st.attempted
    for k in sums:
        assert np.array_equal(sums2[k], sums[k]), k


@pytest.fixture(scope="module")
def cfg0_plain_run_nodes(gpu_ctx):
    """A large plain run on configuration 0's site (FP32 tracer): node counts
    over R realizations."""
    s = sites.config0()
    rng = np.random.default_rng(98)
    R = 60000
    d = draw_realizations(s.wells, s.c_dist, s.p_dist, s.t_dist, R, rng)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"], moments=False)
    counts, _ = trace_counts(s, d, coef, _lib.FP32)
    return s, counts, R


def test_node_control_variate_is_unbiased_over_the_lattice(cfg0_plain_run_nodes):
    """control_variate_nodes on configuration 0's site (8,192 realizations to
    train, 16,384 to correct, 2^17 surrogate evaluations): every node's
    estimate is consistent with a 60,000-realization plain run (z-scores over
    the nodes whose probability lies between 0.02 and 0.98, and the largest
    difference over all nodes), the field outside the window is the
    plain estimate of all traced realizations, and the window nodes'
    variance is below that of the plain estimate from the same traced
    realizations."""
    s, counts, R = cfg0_plain_run_nodes
    res = control_variate_nodes(s, n_train=8192, n_correct=16384, n_surrogate=1 << 17,
                                train_steps=1500, batch=512, lr=1e-3, seed=21, n_hidden=256,
                                qmc_replicates=16)
    # the overlapped and the sequential pipeline give the same estimate, up
    # to the order of the surrogate sums' FP64 atomic adds (amplified by the
    # coefficients of nearly flat cells), far below the standard errors
    seq = control_variate_nodes(s, n_train=8192, n_correct=16384, n_surrogate=1 << 17,
                                train_steps=1500, batch=512, lr=1e-3, seed=21, n_hidden=256,
                                qmc_replicates=16, overlap=False)
    np.testing.assert_allclose(seq.p, res.p, rtol=0, atol=1e-6)
    np.testing.assert_allclose(seq.var, res.var, rtol=1e-3, atol=1e-12)
    p_ref = counts / R
    var_ref = p_ref * (1 - p_ref) / R
    assert res.p.shape == counts.shape
    # z-scores where both estimates are near normal (p between 0.02 and 0.98)
    live = (p_ref > 0.02) & (p_ref < 0.98)
    z = (res.p - p_ref)[live] / np.sqrt(res.var + var_ref)[live]
    inside = res.inside & live
    gain = res.var_plain[res.inside].sum() / res.var[res.inside].sum()
    worst = res.var_plain.max() / res.var.max()
    print(f"nodes {live.sum()} ({inside.sum()} in the window), max |z| {np.abs(z).max():.2f}, "
          f"rms z {np.sqrt(np.mean(z ** 2)):.2f}, window gain {gain:.2f}, worst-node ratio "
          f"{worst:.2f}, outside mass {res.outside_mass:.3g} of {res.total_mass:.3g}, "
          f"max |p - p_ref| {np.abs(res.p - p_ref).max():.4f}", flush=True)
    assert inside.sum() > 150
    assert np.abs(z).max() < 6.0
    assert np.sqrt(np.mean(z ** 2)) < 1.5
    assert np.abs(res.p - p_ref).max() < 0.03
    assert np.array_equal(res.p[~res.inside], res.p_plain[~res.inside])
    assert res.outside_mass <= 0.01 * res.total_mass
    assert gain > 1.5, gain
