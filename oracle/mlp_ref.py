"""FP64 CPU restatement of the learned indicator's network (forward and
backward), the checker of the GPU training and inference kernels.

TEST INFRASTRUCTURE ONLY: only tests/ and ``__graft_entry__.smoke()`` may
import it.  The network is the library's own (the reference has no learned
indicator), so its checks are mathematical: tests/test_mlp_ref.py pins the
gradient against central finite differences.

Parameter layout (canonical, 8705 values):
    W1[64][4], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4
W[j][k] is the weight from input k to unit j.  Hidden activation: GELU in
its tanh form, 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))); output:
sigmoid.  Loss: 1/(2n) sum (y - t)^2.
"""
from __future__ import annotations

import numpy as np

H, IN = 64, 4
NP = H * IN + H + H * H + H + H * H + H + H + 1
_C = np.sqrt(2.0 / np.pi)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H * H].reshape(H, H); o += H * H
    b3 = p[o:o + H]; o += H
    w4 = p[o:o + H]; o += H
    b4 = p[o]
    return W1, b1, W2, b2, W3, b3, w4, b4


def pack(W1, b1, W2, b2, W3, b3, w4, b4):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3, w4, [b4]])


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(_C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(_C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * _C * (1.0 + 3 * 0.044715 * x * x)


def forward(p, X):
    """X [n, 4] -> y [n] (and the activations for backward)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1.T + b1
    a1 = gelu(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    z3 = a2 @ W3.T + b3
    a3 = gelu(z3)
    z4 = a3 @ w4 + b4
    y = 1.0 / (1.0 + np.exp(-z4))
    return y, (X, z1, a1, z2, a2, z3, a3, y)


def loss_grad(p, X, t):
    """Loss 1/(2n) sum (y - t)^2 and its gradient (canonical layout)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p)
    y, (X, z1, a1, z2, a2, z3, a3, _) = forward(p, X)
    n = X.shape[0]
    r = y - np.asarray(t, dtype=np.float64)
    loss = 0.5 * (r * r).sum() / n
    d4 = r * y * (1.0 - y) / n                     # [n]
    gb4 = d4.sum()
    gw4 = a3.T @ d4
    d3 = (d4[:, None] * w4[None, :]) * gelu_grad(z3)  # [n, H]
    gW3 = d3.T @ a2
    gb3 = d3.sum(0)
    d2 = (d3 @ W3) * gelu_grad(z2)
    gW2 = d2.T @ a1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * gelu_grad(z1)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3, gw4, gb4)


def feature(f):
    """Network input of a modal energy fraction f: 0.1 ln(f + 1e-12)."""
    return 0.1 * np.log(np.asarray(f, dtype=np.float64) + 1e-12)
