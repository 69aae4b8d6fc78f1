"""FP64 restatement of the cost-model MLP (TEST INFRASTRUCTURE ONLY).

Forward, mean-squared-error loss and parameter gradients of the F -> 64 -> 64
-> 1 ReLU network in the parameter layout of include/rf_mlp.h, in NumPy
float64, for checking the HIP kernels. Pinned itself by central finite
differences (tests/test_mlp.py)."""
import numpy as np

H = 64


def unpack(p, F):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, x, F):
    W1, b1, W2, b2, W3, b3 = unpack(p, F)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3 + b3


def loss(p, x, t, F):
    y = forward(p, x, F)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def gradients(p, x, t, F):
    W1, b1, W2, b2, W3, b3 = unpack(p, F)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = len(x)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3 + b3
    dy = 2.0 * (y - t) / n
    dW3 = h2.T @ dy
    db3 = dy.sum()
    dz2 = np.outer(dy, W3) * (z2 > 0)
    dW2 = dz2.T @ h1
    db2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    dW1 = dz1.T @ x
    db1 = dz1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]])
