"""Pins the CPU reference of the ELP network: its fp32 gradient against
central finite differences, and its bf16 rounding against exact bfloat16
values."""
import numpy as np
import pytest

import mlp_ref as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-3, 65504.0], np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)     # above the tie: rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) / 3.0e-3 < 2 ** -8
    assert r[5] == 65536.0


def _loss64(p, X, T):
    """float64 forward of the same network, written independently"""
    p = np.asarray(p, np.float64)
    W1 = p[M.OFF_W1:M.OFF_B1].reshape(16, 64)
    b1 = p[M.OFF_B1:M.OFF_W2]
    W2 = p[M.OFF_W2:M.OFF_B2].reshape(64, 64)
    b2 = p[M.OFF_B2:M.OFF_W3]
    W3 = p[M.OFF_W3:M.OFF_B3].reshape(64, 64)
    b3 = p[M.OFF_B3:M.OFF_W4]
    w4 = p[M.OFF_W4:M.OFF_B4]
    h = np.maximum(np.asarray(X, np.float64) @ W1 + b1, 0)
    h = np.maximum(h @ W2 + b2, 0)
    h = np.maximum(h @ W3 + b3, 0)
    y = h @ w4 + p[M.OFF_B4]
    return np.mean((y - T) ** 2)


def test_fp32_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = (rng.standard_normal(M.NPARAM) * 0.2).astype(np.float32)
    X = rng.standard_normal((37, 16)).astype(np.float32)
    T = rng.standard_normal(37).astype(np.float32)
    _, loss, g = M.forward_backward(p, X, T, bf16_ops=False)
    assert loss == pytest.approx(_loss64(p, X, T), rel=1e-5)
    idx = rng.choice(M.NPARAM, 80, replace=False)
    for i in idx:
        h = 1e-5
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[i] += h
        pm[i] -= h
        fd = (_loss64(pp, X, T) - _loss64(pm, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 + 1e-3 * abs(fd), (i, fd, g[i])


def test_transposed_weight_is_detected():
    # a transposed W2 (a plausible indexing mistake) changes the output
    rng = np.random.default_rng(1)
    p = (rng.standard_normal(M.NPARAM) * 0.3).astype(np.float32)
    q = p.copy()
    q[M.OFF_W2:M.OFF_B2] = p[M.OFF_W2:M.OFF_B2].reshape(64, 64).T.ravel()
    X = rng.standard_normal((16, 16)).astype(np.float32)
    assert np.abs(M.forward_backward(p, X) - M.forward_backward(q, X)).max() > 1e-2
