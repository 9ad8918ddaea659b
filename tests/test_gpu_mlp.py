"""The surrogate network's BF16 MFMA kernels against the CPU oracle, and the
learned control variate against the plain estimator."""
import numpy as np
import pytest

from oracle import mlp_ref
from onekapy_amd import sites, _lib
from onekapy_amd.surrogate import (SurrogateMLP, cell_probabilities, choose_window,
                                   control_variate, trace_counts)
from onekapy_amd.stochastic import draw_realizations, fit_and_draw

pytestmark = pytest.mark.gpu


def _net_and_batch(nin=34, nh=128, nout=256, B=128, seed=0):
    net = SurrogateMLP(nin, nh, nout, max_batch=256, seed=seed)
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (B, nin)).astype(np.float32)
    marks = rng.integers(0, 26, (B, nout)).astype(np.uint8)
    return net, x, marks


def test_forward_matches_oracle(gpu_ctx):
    net, x, marks = _net_and_batch()
    probs = net.predict(x)
    p = {k: v.astype(np.float64) for k, v in net.split(net.get_params()).items()}
    _, _, ref = mlp_ref.loss_and_grad(p, x, marks / 25.0, bf16=True)
    assert np.abs(probs - ref).max() < 2e-3


def test_accumulate_sums_match_the_predictions(gpu_ctx):
    """oka_mlp_accumulate's column sums (s, s^2, d = y - s, d^2) against the
    network's own probabilities summed on the host: two passes of max_batch
    rows, the second ragged (rows beyond the input count nothing)."""
    net, _, _ = _net_and_batch(seed=4)
    rng = np.random.default_rng(4)
    n = 300
    x = rng.normal(0, 1, (n, 34)).astype(np.float32)
    marks = rng.integers(0, 26, (n, 256)).astype(np.uint8)
    acc = net.accumulate(x, marks, 1.0 / 25.0)
    p = net.predict(x).astype(np.float64)
    d = marks / np.float32(25.0) - p
    for key, ref in (("s", p.sum(0)), ("s2", (p * p).sum(0)), ("d", d.sum(0)),
                     ("d2", (d * d).sum(0))):
        np.testing.assert_allclose(acc[key], ref, rtol=1e-5, atol=1e-5, err_msg=key)


def test_gradient_matches_oracle(gpu_ctx):
    net, x, marks = _net_and_batch(seed=3)
    g, loss = net.grad(x, marks, 1.0 / 25.0)
    p = {k: v.astype(np.float64) for k, v in net.split(net.get_params()).items()}
    ref_loss, ref, _ = mlp_ref.loss_and_grad(p, x, marks / 25.0, bf16=True)
    assert abs(loss - ref_loss) < 1e-4 * abs(ref_loss) + 1e-6
    gs = net.split(g)
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        a, b = gs[name].astype(np.float64).ravel(), ref[name].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 2e-2, (name, rel)
        assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.9995, name


def test_full_size_products_match_oracle(gpu_ctx):
    """At the control variate's own shape (34 -> 512 -> 512 -> 4096, batch
    1024) the large products run on 128 x 128 block tiles (the small tests
    above use 64 x 64): forward probabilities, loss and every gradient
    against the CPU oracle, with the same bounds."""
    net = SurrogateMLP(34, 512, 4096, max_batch=1024, seed=11)
    rng = np.random.default_rng(12)
    x = rng.normal(0, 1, (1024, 34)).astype(np.float32)
    marks = rng.integers(0, 26, (1024, 4096)).astype(np.uint8)
    p = {k: v.astype(np.float64) for k, v in net.split(net.get_params()).items()}
    ref_loss, ref, ref_probs = mlp_ref.loss_and_grad(p, x, marks / 25.0, bf16=True)
    assert np.abs(net.predict(x) - ref_probs).max() < 2e-3
    g, loss = net.grad(x, marks, 1.0 / 25.0)
    assert abs(loss - ref_loss) < 1e-4 * abs(ref_loss) + 1e-6
    gs = net.split(g)
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        a, b = gs[name].astype(np.float64).ravel(), ref[name].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 2e-2, (name, rel)
        assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.9995, name


@pytest.mark.parametrize("nh,nout,B", [(512, 4096, 1024), (128, 256, 192)],
                         ids=["cfg3_network_b1024", "small_b192"])
def test_training_step_gradient_matches_oracle(gpu_ctx, nh, nout, B):
    """The captured training step computes the weight and bias gradients in
    split-K slices that the Adam kernel sums (at batch 1,024 on the 512-wide
    network: dW3 in 2, dW2 in 4, dW1 in 8; at batch 192, a multiple of 64 but
    not of 128, no dW splits and only layer 3's dX does, in 2, on 64 x 64
    tiles; the gradient call splits no dW).
    With eps far above every gradient, one Adam step moves each parameter by
    -lr g / eps (m = g, v-hat = g^2 at step 1), so the step's own gradient
    can be read back and checked against the CPU oracle like the gradient
    call's."""
    nin = 34
    net = SurrogateMLP(nin, nh, nout, max_batch=B, seed=21)
    rng = np.random.default_rng(22)
    x = rng.normal(0, 1, (B, nin)).astype(np.float32)
    marks = rng.integers(0, 26, (B, nout)).astype(np.uint8)
    p0 = net.get_params().astype(np.float64)
    order = np.arange(B, dtype=np.int32)
    losses = np.zeros(1, dtype=np.float32)
    eps, lr = 1e3, 1e3     # the step is g / (1 + |g| / eps): g itself, to 1e-3
    net.ctx.check(net.ctx.lib.oka_mlp_train(
        net.h, B, x.ctypes.data, marks.ctypes.data, 1.0 / 25.0, order.ctypes.data, B, 1, lr,
        0.9, 0.999, eps, losses.ctypes.data, 0))
    step_grad = (p0 - net.get_params().astype(np.float64)) * eps / lr
    p = {k: v for k, v in net.split(p0).items()}
    _, ref, _ = mlp_ref.loss_and_grad(p, x, marks / 25.0, bf16=True)
    gs = net.split(step_grad)
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        a, b = gs[name].ravel(), ref[name].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 2e-2, (name, rel)


def test_training_fits_a_smooth_target(gpu_ctx):
    rng = np.random.default_rng(5)
    nin, nout = 34, 256
    x = rng.normal(0, 1, (8192, nin)).astype(np.float32)
    w = rng.normal(0, 1, (nin, nout))
    y = 1.0 / (1.0 + np.exp(-(x @ w) / np.sqrt(nin) * 3.0))
    marks = np.rint(y * 25.0).astype(np.uint8)
    net = SurrogateMLP(nin, 256, nout, max_batch=512, seed=1)
    losses = net.train(x, marks, 1.0 / 25.0, epochs=8, batch=256, lr=2e-3, seed=2)
    assert losses[-20:].mean() < 0.6 * losses[:5].mean()
    pred = net.predict(x[:512])
    assert np.abs(pred - marks[:512] / 25.0).mean() < 0.08


@pytest.fixture(scope="module")
def cfg0_plain_run(gpu_ctx):
    """A large plain run on configuration 0's site (FP32 tracer): node counts
    over R realizations."""
    s = sites.config0()
    rng = np.random.default_rng(99)
    R = 40000
    d = draw_realizations(s.wells, s.c_dist, s.p_dist, s.t_dist, R, rng)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"])
    counts, _ = trace_counts(s, d, coef, _lib.FP32)
    return s, counts, R


@pytest.mark.parametrize("mode", ["plain", "antithetic", "qmc"])
def test_control_variate_is_unbiased_and_reduces_variance(cfg0_plain_run, mode):
    s, counts, R = cfg0_plain_run
    res = control_variate(s, n_traced=16384, n_surrogate=131072, folds=2, train_steps=1500,
                          batch=512, lr=1e-3, seed=7, n_hidden=256,
                          antithetic=mode == "antithetic", qmc_replicates=16 if mode == "qmc" else 0,
                          qmc_pow2=True)
    # the reference: the large plain run over the same window
    p_ref = cell_probabilities(counts, R, res.window)
    var_ref = p_ref * (1 - p_ref) / R
    z = (res.p_cv - p_ref) / np.sqrt(res.var_cv + var_ref + 1e-12)
    live = res.var_cv + var_ref > 1e-9
    assert np.abs(z[live]).max() < 6.0
    assert np.sqrt(np.mean(z[live] ** 2)) < 1.5
    gain = res.var_plain.sum() / res.var_cv.sum()
    assert gain > 3.0, gain


@pytest.fixture(scope="module")
def cfg1_fp64_plain_run(gpu_ctx):
    """A large plain FP64-exact run on configuration 1's site (25 wells, 1000
    particles, 20 years, 400^2 nodes): node counts over R realizations."""
    s = sites.config1()
    rng = np.random.default_rng(1234)
    R = 100000
    d = draw_realizations(s.wells, s.c_dist, s.p_dist, s.t_dist, R, rng)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"], moments=False)
    counts, _ = trace_counts(s, d, coef, _lib.FP64_EXACT)
    return s, counts, R


def test_control_variate_config3_network_on_config1_site(cfg1_fp64_plain_run):
    """The learned control variate at configuration 3's network shape (34 ->
    512 -> 512 -> 64 x 64 cell logits) on configuration 1's site, at reduced
    sizes (32,768 traced realizations, 2^19 scrambled Sobol' surrogate
    evaluations): unbiased against a 10^5-realization plain FP64-exact run
    (per-cell z-scores) and with less variance than the plain estimate from
    the same traced realizations."""
    s, counts, R = cfg1_fp64_plain_run
    res = control_variate(s, n_traced=32768, n_surrogate=1 << 19, folds=2, train_steps=1200,
                          batch=2048, lr=5e-3, seed=31, n_hidden=512,
                          precision=_lib.FP64_EXACT, qmc_replicates=64, qmc_pow2=True)
    assert (res.window.nrows, res.window.ncols) == (64, 64)
    p_ref = cell_probabilities(counts, R, res.window)
    var_ref = p_ref * (1 - p_ref) / R
    z = (res.p_cv - p_ref) / np.sqrt(res.var_cv + var_ref + 1e-12)
    live = res.var_cv + var_ref > 1e-9
    gain = res.var_plain.sum() / res.var_cv.sum()
    worst = res.var_plain.max() / res.var_cv[res.var_plain.argmax()]
    print(f"cfg3 network on cfg1: max |z| {np.abs(z[live]).max():.2f}, rms z "
          f"{np.sqrt(np.mean(z[live] ** 2)):.2f}, gain {gain:.1f}, at the worst plain cell "
          f"{worst:.1f}, max |p_cv - p_ref| {np.abs(res.p_cv - p_ref).max():.4f}", flush=True)
    assert live.sum() > 500
    assert np.abs(z[live]).max() < 6.0
    assert np.sqrt(np.mean(z[live] ** 2)) < 1.5
    assert gain > 5.0, gain


@pytest.mark.parametrize("mode", ["plain", "antithetic", "qmc"])
def test_surrogate_sums_compose_draws_fit_features_and_network(gpu_ctx, mode):
    """oka_cv_surrogate_sums against its parts called one by one: the device
    draws, the fit + MVN draw, the standardized features and the network's
    probabilities, summed on the host (squared pair means when antithetic)."""
    from onekapy_amd.stochastic import draw_realizations_device
    from onekapy_amd.surrogate import Standardizer, features
    s = sites.config0()
    net = SurrogateMLP(len(s.wells) + 9, 128, 128, max_batch=256, seed=3)
    seed, first, n = 0xBEEF, 4096, 1000   # four network passes, the last one ragged
    antithetic, qmc = mode == "antithetic", mode == "qmc"
    d = draw_realizations_device(s.wells, s.c_dist, s.p_dist, s.t_dist, n, seed, first,
                                 antithetic, qmc)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"], moments=False)
    feats = features(d, coef)
    std = Standardizer(feats)
    prob = net.predict(std(feats)).astype(np.float64)
    acc = net.surrogate_sums(s, n, seed, first, std, antithetic, qmc)
    # FP32 partial sums per lane, FP64 across workgroups
    np.testing.assert_allclose(acc["s"], prob.sum(0), rtol=1e-5, atol=1e-6)
    if antithetic:
        pm = 0.5 * (prob[0::2] + prob[1::2])
        np.testing.assert_allclose(acc["s2"], (pm * pm).sum(0), rtol=1e-5, atol=1e-6)
        # reflected pairs are negatively correlated over the outputs as a whole:
        # var(pair mean) < var(s) / 2, i.e. sum pm^2 < (sum s^2 + (sum s)^2 / n) / 4
        assert (pm * pm).sum() < 0.25 * ((prob * prob).sum() + (prob.sum(0) ** 2).sum() / n)
    else:
        np.testing.assert_allclose(acc["s2"], (prob * prob).sum(0), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("npts,max_batch", [(1000, 256), (2048, 1024), (3072, 2048)],
                         ids=["per_replicate", "batched", "batched_straddling_passes"])
def test_qmc_replicates_in_one_call_match_separate_calls(gpu_ctx, npts, max_batch):
    """oka_cv_surrogate_sums_qmc: row k equals oka_cv_surrogate_sums with
    OKA_QMC and seed k, up to the order of the workgroups' FP32 LDS and FP64
    atomic additions.  Point counts that are multiples of 1024 run all
    replicates as one sequence, whose network passes cover several replicates
    (3 x 3072 points in passes of 2048 rows)."""
    from onekapy_amd.stochastic import draw_realizations_device
    from onekapy_amd.surrogate import Standardizer, _surrogate_sums_qmc, features
    s = sites.config0()
    net = SurrogateMLP(len(s.wells) + 9, 128, 128, max_batch=max_batch, seed=5)
    d = draw_realizations_device(s.wells, s.c_dist, s.p_dist, s.t_dist, 512, 3)
    _, _, coef = fit_and_draw(s.wells_xy(), s.xt, s.yt, s.base, s.observations, d["K"], d["b"],
                              d["q"], d["z"], moments=False)
    std = Standardizer(features(d, coef))
    seeds = [11, 2 ** 63 + 5, 77]
    rows = _surrogate_sums_qmc(net, s, npts, seeds, std)
    for k, sd in enumerate(seeds):
        one = net.surrogate_sums(s, npts, sd, 0, std, qmc=True)["s"]
        np.testing.assert_allclose(rows[k], one, rtol=1e-6, atol=1e-9)
    assert not np.allclose(rows[0], rows[1])


def test_replayed_training_graph_matches_one_step_calls(gpu_ctx):
    """oka_mlp_train captures its steps 16 at a time as a HIP graph and replays
    it.  Forty steps in one call (two replays of the 16-step graph, one of an
    8-step graph) must leave the parameters bit-identical to forty one-step
    calls (each its own graph, launched once) with other network work on the
    device between them.  Gradients are reduced without atomics, so the only
    difference allowed is the loss's per-workgroup atomic order."""
    rng = np.random.default_rng(9)
    nin, nout, batch, nsteps = 34, 256, 128, 40
    x = rng.normal(0, 1, (1024, nin)).astype(np.float32)
    marks = rng.integers(0, 26, (1024, nout)).astype(np.uint8)
    order = rng.integers(0, 1024, nsteps * batch).astype(np.int32)
    nets = [SurrogateMLP(nin, 128, nout, max_batch=256, seed=4) for _ in range(2)]
    assert np.array_equal(nets[0].get_params(), nets[1].get_params())

    def train(net, first, k):
        losses = np.zeros(k, dtype=np.float32)
        o = np.ascontiguousarray(order[first * batch:(first + k) * batch])
        net.ctx.check(net.ctx.lib.oka_mlp_train(
            net.h, x.shape[0], x.ctypes.data, marks.ctypes.data, 1.0 / 25.0, o.ctypes.data,
            batch, k, 2e-3, 0.9, 0.999, 1e-8, losses.ctypes.data, 0))
        return losses

    whole = train(nets[0], 0, nsteps)
    steps = []
    for i in range(nsteps):
        steps.append(train(nets[1], i, 1)[0])
        nets[1].predict(x[:64])            # eager forward passes between the graphs
        nets[1].accumulate(x[:128], marks[:128], 1.0 / 25.0)
    assert np.array_equal(nets[0].get_params(), nets[1].get_params())
    np.testing.assert_allclose(whole, np.array(steps), rtol=1e-5)
    assert whole[-8:].mean() < whole[:8].mean()
