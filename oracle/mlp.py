"""FP64 CPU oracle of the surrogate MLP (TEST INFRASTRUCTURE ONLY).

The reference has no learned component, so this restates the network the
library defines (lqt.h, lqt_mlp_*): x -> GELU(W1 x + b1) -> GELU(W2 h + b2)
-> W3 h + b3; loss = mean over rows of sum_j mask_j (o_j - y_j)^2 on the
regression outputs plus the masked binary cross-entropy with logits on the
last outputs.  Gradients are written out by hand (chain rule) and pinned in
tests/test_oracle_mlp.py by central finite differences.

`rounding=True` rounds to bfloat16 (round to nearest even) where the GPU
stores bfloat16: inputs, weights, hidden activations and the backpropagated
signals, so the GPU gradients can be compared tightly.
"""
from __future__ import annotations

import math

import numpy as np

try:
    from scipy.special import erf as _erf
except Exception:  # pragma: no cover
    _erf = np.vectorize(math.erf)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def forward(p, x, rounding=False):
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    w1, b1, w2, b2, w3, b3 = [np.asarray(a, np.float64) for a in p]
    X = r(x)
    a1 = X @ r(w1).T + b1
    h1 = r(gelu(a1))
    a2 = h1 @ r(w2).T + b2
    h2 = r(gelu(a2))
    out = h2 @ r(w3).T + b3
    return dict(X=X, a1=a1, h1=h1, a2=a2, h2=h2, out=out)


def loss_and_grads(p, x, y, mask=None, n_reg=4, n_bce=1, rounding=False):
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    f = forward(p, x, rounding)
    o = f["out"]
    y = np.asarray(y, np.float64)
    M = o.shape[0]
    w = np.ones_like(o) if mask is None else np.asarray(mask, np.float64)
    L = 0.0
    g = np.zeros_like(o)
    d = o[:, :n_reg] - y[:, :n_reg]
    L += np.sum(w[:, :n_reg] * d * d)
    g[:, :n_reg] = 2.0 * w[:, :n_reg] * d
    ob, tb, wb = o[:, n_reg:n_reg + n_bce], y[:, n_reg:n_reg + n_bce], w[:, n_reg:n_reg + n_bce]
    L += np.sum(wb * (np.maximum(ob, 0) - ob * tb + np.log1p(np.exp(-np.abs(ob)))))
    g[:, n_reg:n_reg + n_bce] = wb * (1.0 / (1.0 + np.exp(-ob)) - tb)
    L /= M
    g = r(g / M)
    w1, b1, w2, b2, w3, b3 = [np.asarray(a, np.float64) for a in p]
    gw3 = g.T @ f["h2"]
    gb3 = g.sum(0)
    g2 = r((g @ r(w3)) * gelu_grad(f["a2"]))
    gw2 = g2.T @ f["h1"]
    gb2 = g2.sum(0)
    g1 = r((g2 @ r(w2)) * gelu_grad(f["a1"]))
    gw1 = g1.T @ f["X"]
    gb1 = g1.sum(0)
    return L, [gw1, gb1, gw2, gb2, gw3, gb3]


def init(n_in=6, hidden=128, n_out=5, seed=0):
    rng = np.random.default_rng(seed)
    return [rng.normal(0, math.sqrt(2.0 / n_in), (hidden, n_in)), rng.normal(0, 0.1, hidden),
            rng.normal(0, math.sqrt(2.0 / hidden), (hidden, hidden)), rng.normal(0, 0.1, hidden),
            rng.normal(0, math.sqrt(1.0 / hidden), (n_out, hidden)), rng.normal(0, 0.1, n_out)]
