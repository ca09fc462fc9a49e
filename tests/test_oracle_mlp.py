"""Pins of the MLP oracle: hand-written gradients against central finite
differences, and the bfloat16 rounding against known bit patterns."""
import numpy as np

from oracle import mlp as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-45], np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert np.isfinite(r[4])
    assert r[5] == 0.0 or abs(r[5]) < 1e-38


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init(6, 16, 5, seed=1)
    x = rng.normal(size=(9, 6))
    y = rng.normal(size=(9, 5))
    y[:, 4] = rng.integers(0, 2, 9)
    mask = (rng.random((9, 5)) > 0.2).astype(float)
    L, g = M.loss_and_grads(p, x, y, mask)
    eps = 1e-6
    for k, P in enumerate(p):
        flat = P.reshape(-1)
        for idx in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = M.loss_and_grads(p, x, y, mask)
            flat[idx] = old - eps
            lm, _ = M.loss_and_grads(p, x, y, mask)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = g[k].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, an)


def test_gelu_derivative():
    x = np.linspace(-5, 5, 101)
    fd = (M.gelu(x + 1e-6) - M.gelu(x - 1e-6)) / 2e-6
    assert np.allclose(M.gelu_grad(x), fd, atol=1e-8)
