"""The node-resolution control-variate estimator (surrogate.node_estimate) on
synthetic data whose node probabilities are known exactly: no GPU, only the
estimator's algebra from per-part sums.

Each synthetic realization has a latent u ~ U(0, 1); cell c's surrogate value
is s_c = clip(u + a_c + noise, 0, 1) and node k of the cell is visited when
u < t_k, so p_k = t_k exactly and y_k is correlated with s_c and with the
neighbouring cells' values.  The surrogate means come from independent
replicate samples, as the QMC replicates do."""
import numpy as np

from onekapy_amd.surrogate import PAIRS, _node_image, _window_slices, neighbours, node_estimate


class _Win:
    def __init__(self, row0, col0, cell, nrows, ncols):
        self.row0, self.col0, self.cell, self.nrows, self.ncols = row0, col0, cell, nrows, ncols


WIN = _Win(0, 0, 2, 3, 4)              # 3 x 4 cells of 2 x 2 nodes
NCELL, NK = 12, 4
THRESH = np.linspace(0.05, 0.9, NCELL * NK).reshape(NCELL, NK)
OFFSET = np.linspace(-0.2, 0.2, NCELL)


def _sample(rng, n):
    """s [n, cells] and y [n, cells, nk]: node k of cell a is visited when
    u < t; each cell's surrogate value is a noisy ramp in u."""
    u = rng.random(n)
    s = np.clip(u[:, None] + OFFSET[None, :] + rng.normal(0, 0.05, (n, NCELL)), 0.0, 1.0)
    y = (u[:, None, None] < THRESH[None, :, :]).astype(np.float64)
    return s, y


def _sums(s, y, parts):
    """The sums oka_cv_node_sums returns, from per-realization values."""
    nb = neighbours(WIN)
    nr, nc = WIN.nrows, WIN.ncols
    n = s.shape[0]
    b = [n * p // parts for p in range(parts + 1)]
    pair = np.full((NCELL, len(PAIRS)), -1)
    for a in range(NCELL):
        r, c = divmod(a, nc)
        for h, (dr, dc) in enumerate(PAIRS):
            if r + dr < nr and 0 <= c + dc < nc:
                pair[a, h] = (r + dr) * nc + c + dc
    sp = np.concatenate([s, np.zeros((n, 1))], axis=1)          # index -1: an absent cell
    out = {k: [] for k in ("s", "ss", "y", "ys")}
    for p in range(parts):
        ss, yy, spp = s[b[p]:b[p + 1]], y[b[p]:b[p + 1]], sp[b[p]:b[p + 1]]
        out["s"].append(ss.sum(0))
        out["ss"].append((ss[:, :, None] * spp[:, pair]).sum(0))
        out["y"].append(yy.sum(0))
        out["ys"].append((yy[:, :, :, None] * spp[:, nb][:, :, None, :]).sum(0))
    return {k: np.array(v) for k, v in out.items()}, np.diff(b)


def _reps(rng, n, reps):
    return np.array([_sample(rng, n)[0].mean(0) for _ in range(reps)])


def test_node_estimate_is_unbiased_and_its_variance_is_right():
    rng = np.random.default_rng(11)
    nb = neighbours(WIN)
    est, var = [], []
    for _ in range(1200):
        s, y = _sample(rng, 300)
        sums, n_rows = _sums(s, y, 2)
        p, v, beta = node_estimate(sums, n_rows, nb, _reps(rng, 500, 8), min_events=0)
        est.append(p)
        var.append(v)
    est, var = np.array(est), np.array(var)
    # unbiased: the mean over repetitions is within 4.5 standard errors of t
    se = est.std(0, ddof=1) / np.sqrt(len(est))
    z = (est.mean(0) - THRESH) / se
    assert np.abs(z).max() < 4.5, np.abs(z).max()
    # the variance estimate matches the spread of the estimates (to 15 %)
    ratio = var.mean(0) / est.var(0, ddof=1)
    assert 0.85 < ratio.min() and ratio.max() < 1.15, (ratio.min(), ratio.max())
    # and it is below the plain estimate's t (1 - t) / n at every node, by a
    # third on average (a step in u against a few ramps in u)
    plain = THRESH * (1 - THRESH) / 300
    assert (var.mean(0) < plain).all(), (var.mean(0) / plain).max()
    assert (var.mean(0) / plain).mean() < 0.67


def test_node_estimate_without_correlation_is_the_plain_mean():
    """With s constant over the realizations the coefficients are 0 and the
    estimate is each node's plain mean."""
    rng = np.random.default_rng(3)
    _, y = _sample(rng, 300)
    s = np.full((300, NCELL), 0.25)
    sums, n_rows = _sums(s, y, 3)
    reps = np.full((4, NCELL), 0.3)
    p, v, beta = node_estimate(sums, n_rows, neighbours(WIN), reps, min_events=0)
    assert np.abs(beta).max() < 1e-12
    np.testing.assert_allclose(p, y.mean(0), rtol=0, atol=1e-12)
    ybar = y.reshape(3, 100, NCELL, NK).mean(1)
    np.testing.assert_allclose(v, (ybar * (1 - ybar)).sum(0) / 9 / 100, rtol=1e-9, atol=1e-15)


def test_neighbours_and_pair_offsets():
    """Neighbour m of cell (r, c) is cell (r + m // 3 - 1, c + m % 3 - 1); the
    13 half-offsets and their mirror images are the 25 offsets of a 5 x 5
    block, each once."""
    nb = neighbours(WIN)
    assert nb.shape == (12, 9)
    assert list(nb[5]) == [0, 1, 2, 4, 5, 6, 8, 9, 10]          # cell (1, 1)
    assert list(nb[0]) == [-1, -1, -1, -1, 0, 1, -1, 4, 5]       # corner (0, 0)
    offs = set(PAIRS) | {(-a, -b) for a, b in PAIRS}
    assert len(PAIRS) == 13 and len(offs) == 25
    assert offs == {(a, b) for a in range(-2, 3) for b in range(-2, 3)}


def test_node_image_and_window_slices():
    """Cell (r, c) node (i, j) -> window node (r cell + i, c cell + j); the
    window's part on the grid, clipped at the grid's edges."""
    w = _Win(-2, 5, 3, 2, 4)
    a = np.arange(2 * 4 * 9).reshape(8, 9)       # [cells, cell^2]: value = 9 cell + k
    img = _node_image(a, w)
    assert img.shape == (6, 12)
    for r in range(2):
        for c in range(4):
            for i in range(3):
                for j in range(3):
                    assert img[r * 3 + i, c * 3 + j] == 9 * (r * 4 + c) + i * 3 + j
    gr, gc, wr, wc = _window_slices(w, (10, 15))
    assert (gr, gc, wr, wc) == (slice(0, 4), slice(5, 15), slice(2, 6), slice(0, 10))
    assert _window_slices(_Win(20, 0, 2, 3, 3), (10, 10)) is None


def test_rare_nodes_take_no_control():
    """A node visited by fewer than min_events of the other part's
    realizations keeps its plain mean (its coefficients are zero)."""
    rng = np.random.default_rng(8)
    s, y = _sample(rng, 300)
    sums, n_rows = _sums(s, y, 2)
    reps = _reps(rng, 500, 4)
    p, v, beta = node_estimate(sums, n_rows, neighbours(WIN), reps, min_events=10 ** 6)
    assert np.all(beta == 0.0)
    np.testing.assert_allclose(p, y.mean(0), rtol=0, atol=1e-12)
    _, _, b30 = node_estimate(sums, n_rows, neighbours(WIN), reps, min_events=30)
    rare = np.minimum(y[:150].sum(0), 150 - y[:150].sum(0)) < 30   # events of part 0
    assert rare.any() and np.all(np.abs(b30[rare]).sum(-1) >= 0)


def test_surrogate_mean_term_on_a_window_smaller_than_the_pair_offsets():
    """The surrogate means' share of the variance, beta_bar' Cov(mu) beta_bar,
    formed through the half-offset pair sums, equals numpy's covariance of the
    replicate means over each cell's neighbours, on a 2 x 2-cell window whose
    pair offsets reach beyond it (absent neighbours count as 0)."""
    win = _Win(0, 0, 2, 2, 2)
    ncell, nk = 4, 4
    rng = np.random.default_rng(31)
    n = 400
    u = rng.random(n)
    off = np.array([-0.1, 0.0, 0.05, 0.1])
    s = np.clip(u[:, None] + off[None, :] + rng.normal(0, 0.05, (n, ncell)), 0.0, 1.0)
    t = np.linspace(0.2, 0.8, ncell * nk).reshape(ncell, nk)
    y = (u[:, None, None] < t[None, :, :]).astype(np.float64)
    nb = neighbours(win)
    pair = np.full((ncell, len(PAIRS)), -1)
    for a in range(ncell):
        r, c = divmod(a, 2)
        for h, (dr, dc) in enumerate(PAIRS):
            if r + dr < 2 and 0 <= c + dc < 2:
                pair[a, h] = (r + dr) * 2 + c + dc
    sp = np.concatenate([s, np.zeros((n, 1))], axis=1)
    half = n // 2
    sums = {k: [] for k in ("s", "ss", "y", "ys")}
    for sl in (slice(0, half), slice(half, n)):
        sums["s"].append(s[sl].sum(0))
        sums["ss"].append((s[sl][:, :, None] * sp[sl][:, pair]).sum(0))
        sums["y"].append(y[sl].sum(0))
        sums["ys"].append((y[sl][:, :, :, None] * sp[sl][:, nb][:, :, None, :]).sum(0))
    sums = {k: np.array(v) for k, v in sums.items()}
    reps = rng.normal(0.5, 0.02, (12, ncell))
    flat = np.repeat(reps.mean(0, keepdims=True), 12, axis=0)
    p1, v1, b1 = node_estimate(sums, [half, n - half], nb, reps, min_events=0)
    p0, v0, b0 = node_estimate(sums, [half, n - half], nb, flat, min_events=0)
    np.testing.assert_array_equal(b1, b0)
    mu_nb = np.where(nb[None] < 0, 0.0, reps[:, np.where(nb < 0, 0, nb)])   # [R, cells, 9]
    want = np.zeros((ncell, nk))
    for a in range(ncell):
        cov = np.cov(mu_nb[:, a, :], rowvar=False) / reps.shape[0]
        want[a] = np.einsum("km,ml,kl->k", b1[a], cov, b1[a])
    np.testing.assert_allclose(v1 - v0, want, rtol=1e-9, atol=1e-15)
    assert want.max() > 0
