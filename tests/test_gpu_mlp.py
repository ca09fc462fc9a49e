"""Surrogate MLP kernels (BF16 MFMA, FP32 accumulation) against the FP64
oracle: forward outputs, loss and every parameter gradient; Adam training;
and the steered dose sweep against the plain solver sweep."""
import numpy as np
import pytest

from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def _net(lqt, seed=0, hidden=128):
    from lqt1_amd.surrogate import MLP

    m = MLP(6, hidden, 4, 1, seed=seed)
    p = OM.init(6, hidden, 5, seed=seed + 10)
    # weights exactly representable in bf16 so the GPU copies are exact
    p = [OM.bf16(a).astype(np.float32) if a.ndim == 2 else a.astype(np.float32) for a in p]
    m.set_params(p)
    return m, p


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    y = rng.normal(0, 0.5, (n, 5)).astype(np.float32)
    y[:, 4] = rng.integers(0, 2, n)
    mask = (rng.random((n, 5)) > 0.1).astype(np.float32)
    return x, y, mask


def test_forward_matches_oracle(lqt):
    m, p = _net(lqt)
    x, _, _ = _data(1000)
    out = m.forward(x)
    ref = OM.forward([a.astype(np.float64) for a in p], x, rounding=True)["out"]
    err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    assert err < 2e-3, err
    # the whole network in FP64 without bf16 rounding is within bf16 error
    ref64 = OM.forward([a.astype(np.float64) for a in p], x, rounding=False)["out"]
    assert np.max(np.abs(out - ref64)) / np.max(np.abs(ref64)) < 3e-2


def test_gradients_match_oracle(lqt):
    m, p = _net(lqt, 1)
    x, y, mask = _data(256, 1)
    loss, g = m.grads(x, y, mask)
    L, gr = OM.loss_and_grads([a.astype(np.float64) for a in p], x, y, mask, rounding=True)
    assert abs(loss - L) < 1e-3 * max(1.0, abs(L))
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    for k in range(6):
        rel = np.linalg.norm(g[k] - gr[k]) / max(np.linalg.norm(gr[k]), 1e-12)
        print(f"grad {names[k]}: relative L2 error {rel:.2e}")
        assert rel < 2e-2, (names[k], rel)
    _, g64 = OM.loss_and_grads([a.astype(np.float64) for a in p], x, y, mask, rounding=False)
    for k in range(6):
        rel = np.linalg.norm(g[k] - g64[k]) / max(np.linalg.norm(g64[k]), 1e-12)
        assert rel < 8e-2, (names[k], rel)


def test_adam_training_fits_a_smooth_map(lqt):
    from lqt1_amd.surrogate import MLP

    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (8192, 6)).astype(np.float32)
    y = np.zeros((8192, 5), np.float32)
    y[:, 0] = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]
    y[:, 1] = x[:, 3] ** 2 - 0.3
    y[:, 2] = 0.5 * x[:, 4]
    y[:, 3] = np.tanh(x[:, 5] + x[:, 0])
    y[:, 4] = (x[:, 0] + x[:, 1] > 0).astype(np.float32)
    m = MLP(seed=3)
    losses = m.train(x, y, None, epochs=60, batch=512, lr=2e-3, seed=3)
    print("loss first/last epoch:", losses[:16].mean(), losses[-16:].mean())
    assert losses[-16:].mean() < 0.1 * losses[:16].mean()
    out = m.forward(x[:1024])
    assert np.mean(np.abs(out[:, 0] - y[:1024, 0])) < 0.1
    acc = np.mean((out[:, 4] > 0) == (y[:1024, 4] > 0.5))
    assert acc > 0.95
