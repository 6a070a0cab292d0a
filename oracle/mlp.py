"""FP64 NumPy MLP (checker only): forward and backward of the surrogate's
architecture, Linear -> SiLU ... -> Linear, to check the device kernels'
activations and gradients.

TEST INFRASTRUCTURE: imported only by tests/.
"""
from __future__ import annotations

import numpy as np


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(x, params):
    """params: list of (W [out, in], b [out]); SiLU after every layer but the last."""
    acts, pre = [x], []
    h = x
    for i, (W, b) in enumerate(params):
        z = h @ W.T + b
        pre.append(z)
        h = silu(z) if i < len(params) - 1 else z
        acts.append(h)
    return h, acts, pre


def backward(dy, params, acts, pre):
    """Gradients (dW, db) per layer for upstream gradient dy of the output."""
    grads = [None] * len(params)
    g = dy
    for i in range(len(params) - 1, -1, -1):
        W, b = params[i]
        grads[i] = (g.T @ acts[i], g.sum(axis=0))
        if i > 0:
            g = (g @ W) * dsilu(pre[i - 1])
    return grads
