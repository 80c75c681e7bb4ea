"""CPU reference of the learned error-latency profile network (test
infrastructure only): a 16-64-64-64-1 ReLU MLP with the device's rounding
points -- inputs, weights, hidden activations and hidden-layer gradients are
rounded to bfloat16 (round to nearest even) before each matrix product,
products accumulate in float32/float64, biases and the output layer stay in
float32.  Loss: mean squared error over the batch.

`forward_backward(params, X, T, bf16=True)` returns (y, loss, grad) with the
gradient laid out like the device's parameter vector.  With bf16=False it is
the plain fp32 network (for finite-difference checks).
"""
import numpy as np

IN, H = 16, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
NPARAM = OFF_B4 + 1


def bf16(x):
    """round float32 values to bfloat16 (nearest even), returned as float32"""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    return (p[OFF_W1:OFF_B1].reshape(IN, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(H, H), p[OFF_B3:OFF_W4], p[OFF_W4:OFF_B4],
            p[OFF_B4])


def forward_backward(params, X, T=None, bf16_ops=True):
    rd = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float32))
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(params)
    X = np.asarray(X, np.float32)
    B = X.shape[0]
    f64 = np.float64
    Xb = rd(X)
    W1b, W2b, W3b = rd(W1), rd(W2), rd(W3)
    Z1 = (Xb.astype(f64) @ W1b.astype(f64)).astype(np.float32) + b1
    H1 = rd(np.maximum(Z1, 0))
    Z2 = (H1.astype(f64) @ W2b.astype(f64)).astype(np.float32) + b2
    H2 = rd(np.maximum(Z2, 0))
    Z3 = (H2.astype(f64) @ W3b.astype(f64)).astype(np.float32) + b3
    H3 = rd(np.maximum(Z3, 0))
    y = (H3.astype(f64) @ w4.astype(f64)).astype(np.float32) + b4
    if T is None:
        return y
    T = np.asarray(T, np.float32)
    e = y - T
    loss = float(np.mean(e.astype(f64) ** 2))
    dy = (2.0 * e / B).astype(np.float32)
    g = np.zeros(NPARAM, np.float64)
    g[OFF_W4:OFF_B4] = H3.T.astype(f64) @ dy.astype(f64)
    g[OFF_B4] = dy.astype(f64).sum()
    dZ3 = np.where(H3 > 0, np.outer(dy, w4), 0).astype(np.float32)
    g[OFF_B3:OFF_W4] = dZ3.astype(f64).sum(0)
    dZ3b = rd(dZ3)
    g[OFF_W3:OFF_B3] = (H2.T.astype(f64) @ dZ3b.astype(f64)).ravel()
    dH2 = (dZ3b.astype(f64) @ W3b.T.astype(f64)).astype(np.float32)
    dZ2 = np.where(H2 > 0, dH2, 0).astype(np.float32)
    g[OFF_B2:OFF_W3] = dZ2.astype(f64).sum(0)
    dZ2b = rd(dZ2)
    g[OFF_W2:OFF_B2] = (H1.T.astype(f64) @ dZ2b.astype(f64)).ravel()
    dH1 = (dZ2b.astype(f64) @ W2b.T.astype(f64)).astype(np.float32)
    dZ1 = np.where(H1 > 0, dH1, 0).astype(np.float32)
    g[OFF_B1:OFF_W2] = dZ1.astype(f64).sum(0)
    dZ1b = rd(dZ1)
    g[OFF_W1:OFF_B1] = (Xb.T.astype(f64) @ dZ1b.astype(f64)).ravel()
    return y, loss, g.astype(np.float32)
