"""numpy restatement of the learned tile-cost model.  TEST INFRASTRUCTURE
ONLY: imported by tests/ as the checker of the HIP forward, gradient and
Adam kernels (fractal-audio-viz_amd/csrc/sched_mlp.hip); never by the
product.

The network has no counterpart in the reference.  It maps 65 inputs (an
8x8 patch of a preview's log(1 + work) plus log10(zoom) / 100) through
64 and 64 ReLU units to one output.  Weights and inputs are rounded to bf16
(nearest, ties to even) before use; sums and activations are float32 in the
HIP kernels and float64 here.  The loss is mean((y - t)^2); the gradient is
taken with respect to the rounded weights.

Pinning: `gradient` is checked against central finite differences of
`forward` in float64 (tests/test_sched_cpu.py), which fixes every term of
the back-propagation; `adam_step` against the closed form of one Adam step
from zero moments, -lr * sign(g) (up to eps).
"""
import numpy as np

INPUTS = 65
HIDDEN = 64
W1 = 0
B1 = W1 + INPUTS * HIDDEN
W2 = B1 + HIDDEN
B2 = W2 + HIDDEN * HIDDEN
W3 = B2 + HIDDEN
B3 = W3 + HIDDEN
PARAMS = B3 + 1


def bf16_round(values) -> np.ndarray:
    """float32 values rounded to the nearest bf16 (ties to even); inf and NaN
    kept."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    bits = v.view(np.uint32).astype(np.uint64)
    special = (bits & 0x7F800000) == 0x7F800000
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = np.where(special, bits.astype(np.uint32), rounded)
    return out.view(np.float32).reshape(v.shape)


def unpack(w):
    """(W1 [65,64], b1, W2 [64,64], b2, W3 [64], b3) views of a flat vector."""
    w = np.asarray(w)
    return (w[W1:B1].reshape(INPUTS, HIDDEN), w[B1:W2], w[W2:B2].reshape(HIDDEN, HIDDEN), w[B2:W3], w[W3:B3],
            w[B3])


def effective(weights, features):
    """The float64 weights and inputs the network computes with."""
    return (bf16_round(weights).astype(np.float64), bf16_round(features).astype(np.float64))


def forward(w, x):
    """Outputs and activations (h1, h2) for float64 weights w and inputs x."""
    w1, b1, w2, b2, w3, b3 = unpack(w)
    h1 = np.maximum(x @ w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2 + b2, 0.0)
    return h2 @ w3 + b3, h1, h2


def loss(w, x, t) -> float:
    y, _, _ = forward(w, x)
    return float(np.mean((y - t) ** 2))


def gradient(w, x, t):
    """d loss / d w (flat, float64) and the loss."""
    y, h1, h2 = forward(w, x)
    n = x.shape[0]
    _, _, w2, _, w3, _ = unpack(w)
    dy = 2.0 * (y - t) / n
    d2 = np.where(h2 > 0.0, dy[:, None] * w3[None, :], 0.0)
    dh1 = np.where(h1 > 0.0, d2 @ w2.T, 0.0)
    g = np.empty(PARAMS, dtype=np.float64)
    g[W1:B1] = (x.T @ dh1).ravel()
    g[B1:W2] = dh1.sum(axis=0)
    g[W2:B2] = (h1.T @ d2).ravel()
    g[B2:W3] = d2.sum(axis=0)
    g[W3:B3] = h2.T @ dy
    g[B3] = dy.sum()
    return g, float(np.mean((y - t) ** 2))


def adam_step(w, g, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (returns new w, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
