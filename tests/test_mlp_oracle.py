"""The surrogate network's CPU oracle pinned by central finite differences and
by its bf16 rounding rule."""
import numpy as np

from oracle import mlp_ref


def _params(rng, nin, nh, nout, scale=0.5):
    return {"W1": rng.normal(0, scale, (nh, nin)), "b1": rng.normal(0, 0.1, nh),
            "W2": rng.normal(0, scale, (nh, nh)), "b2": rng.normal(0, 0.1, nh),
            "W3": rng.normal(0, scale, (nout, nh)), "b3": rng.normal(0, 0.1, nout)}


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng, 5, 7, 4)
    x = rng.normal(0, 1, (6, 5))
    y = rng.uniform(0, 1, (6, 4))
    loss, g, _ = mlp_ref.loss_and_grad(p, x, y, bf16=False)
    h = 1e-6
    for name, arr in p.items():
        flat = arr.reshape(-1)
        for idx in rng.choice(flat.size, min(flat.size, 6), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _, _ = mlp_ref.loss_and_grad(p, x, y, bf16=False)
            flat[idx] = old - h
            lm, _, _ = mlp_ref.loss_and_grad(p, x, y, bf16=False)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name].reshape(-1)[idx]) <= 1e-7 + 1e-5 * abs(fd), (name, idx)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40])
    out = mlp_ref.to_bf16(vals)
    assert out[0] == 1.0
    assert out[1] == 1.0                     # tie -> even (mantissa bit 0)
    assert out[2] == 1.0 + 2 ** -7           # exactly representable
    assert out[3] == 1.0 + 2 ** -6           # tie -> even
    assert out[4] == -2.5
    assert np.isfinite(out[5])
    assert np.isnan(mlp_ref.to_bf16(np.array([np.nan]))[0])
