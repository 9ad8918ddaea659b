"""CPU oracle of the surrogate network -- TEST INFRASTRUCTURE ONLY.

Forward and backward of the 3-layer perceptron with GELU (erf form) and the
mean binary cross-entropy with logits, in float64.  With bf16=True every
matrix-product operand is rounded to bfloat16 (round to nearest even) at the
points where the GPU kernels store BF16: the input, the weights, the hidden
activations and the back-propagated gradients.  With bf16=False it is the
exact gradient, which the tests pin by central finite differences.
"""
import numpy as np
from scipy.special import erf


def to_bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(f)
    out = rounded.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss_and_grad(p, x, y, bf16=True):
    """p: dict W1 b1 W2 b2 W3 b3; x [B, n_in]; y [B, n_out] in [0, 1].
    Returns (loss, dict of gradients, probabilities)."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    B, nout = y.shape
    X0 = r(x)
    W1, W2, W3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    b1, b2, b3 = (np.asarray(p[k], dtype=np.float64) for k in ("b1", "b2", "b3"))
    Z1 = X0 @ W1.T + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ W2.T + b2
    H2 = r(gelu(Z2))
    Z3 = H2 @ W3.T + b3
    loss = np.mean(np.maximum(Z3, 0) - Z3 * y + np.log1p(np.exp(-np.abs(Z3))))
    G3 = r((sigmoid(Z3) - y) / (B * nout))
    g = {"W3": G3.T @ H2, "b3": G3.sum(0)}
    G2 = r((G3 @ W3) * gelu_grad(Z2))
    g["W2"], g["b2"] = G2.T @ H1, G2.sum(0)
    G1 = r((G2 @ W2) * gelu_grad(Z1))
    g["W1"], g["b1"] = G1.T @ X0, G1.sum(0)
    return loss, g, sigmoid(Z3)
