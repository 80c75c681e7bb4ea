"""The ELP network's BF16-MFMA forward and backward kernels against the CPU
reference with the same bf16 rounding points, plus a training run."""
import numpy as np
import pytest

import mlp_ref as M
from zgraph.mlp import MLP

pytestmark = pytest.mark.gpu


def _data(B, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, 16)).astype(np.float32)
    T = (X[:, :4].sum(1) + 0.5 * np.sin(X[:, 5]) + 0.1 * rng.standard_normal(B)).astype(np.float32)
    return X, T


@pytest.mark.parametrize("B", [1, 16, 37, 1000])
def test_forward_matches_reference(B):
    net = MLP(seed=3)
    X, _ = _data(B, 1)
    y = net.forward(X)
    ref = M.forward_backward(net.params(), X)
    assert np.allclose(y, ref, rtol=1e-3, atol=1e-3 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("B", [16, 53, 512])
def test_gradient_matches_reference(B):
    net = MLP(seed=5)
    X, T = _data(B, 2)
    g, loss = net.gradient(X, T)
    _, rloss, rg = M.forward_backward(net.params(), X, T)
    assert loss == pytest.approx(rloss, rel=1e-3)
    for a, b, name in ((M.OFF_W1, M.OFF_B1, "W1"), (M.OFF_B1, M.OFF_W2, "b1"), (M.OFF_W2, M.OFF_B2, "W2"),
                       (M.OFF_B2, M.OFF_W3, "b2"), (M.OFF_W3, M.OFF_B3, "W3"), (M.OFF_B3, M.OFF_W4, "b3"),
                       (M.OFF_W4, M.OFF_B4, "w4"), (M.OFF_B4, M.NPARAM, "b4")):
        d, r = g[a:b], rg[a:b]
        scale = max(np.abs(r).max(), 1e-6)
        err = np.abs(d - r).max() / scale
        print(f"{name}: max rel err {err:.2e}")
        assert err < 2e-2, name
    # fp32 gradient direction agrees closely too
    _, _, g32 = M.forward_backward(net.params(), X, T, bf16_ops=False)
    cos = float(g @ g32 / (np.linalg.norm(g) * np.linalg.norm(g32)))
    assert cos > 0.99


def test_training_reduces_loss():
    net = MLP(seed=9)
    X, T = _data(4096, 3)
    _, l0 = net.gradient(X, T)
    net.train(X, T, steps=800, batch=256, lr=3e-3, seed=11)
    _, l1 = net.gradient(X, T)
    assert l1 < 0.2 * l0, (l0, l1)
