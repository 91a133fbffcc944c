"""ORACLE (test infrastructure only): FP64 numpy restatement of the learned
closure's network and its gradient.

The product runs this network with its own HIP kernels
(nhqg_amd/csrc/closure.hip); tests compare those kernels' outputs and
gradients with the functions here, which are themselves pinned by central
finite differences (tests/test_closure.py).  The closure has no counterpart
in the reference.

Network: x -> tanh(W1 x + b1) -> tanh(W2 h1 + b2) -> W3 h2 + b3, parameters
as one flat array W1 (width x n_in), b1, W2 (width x width), b2,
W3 (n_out x width), b3; loss = mean((y - t)^2) over batch and outputs.
"""

from __future__ import annotations

import numpy as np


def unpack(w, n_in, width, n_out):
    w = np.asarray(w, dtype=np.float64)
    sizes = [width * n_in, width, width * width, width, n_out * width, n_out]
    parts, o = [], 0
    for s in sizes:
        parts.append(w[o:o + s])
        o += s
    W1 = parts[0].reshape(width, n_in)
    W2 = parts[2].reshape(width, width)
    W3 = parts[4].reshape(n_out, width)
    return W1, parts[1], W2, parts[3], W3, parts[5]


def forward(w, x, n_in, width, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(w, n_in, width, n_out)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.tanh(x @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ W3.T + b3, (x, a1, a2)


def loss_and_grad(w, x, t, n_in, width, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(w, n_in, width, n_out)
    y, (x, a1, a2) = forward(w, x, n_in, width, n_out)
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dy = 2.0 * (y - t) / n
    gW3 = dy.T @ a2
    gb3 = dy.sum(0)
    d2 = (dy @ W3) * (1.0 - a2 ** 2)
    gW2 = d2.T @ a1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (1.0 - a1 ** 2)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def init_params(n_in, width, n_out, seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)

    def glorot(fo, fi):
        lim = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-lim, lim, size=(fo, fi))

    return np.concatenate([glorot(width, n_in).ravel(), np.zeros(width),
                           glorot(width, width).ravel(), np.zeros(width),
                           glorot(n_out, width).ravel(), np.zeros(n_out)])
