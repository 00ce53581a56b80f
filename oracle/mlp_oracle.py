"""CPU restatement of the lambda* model (test infrastructure only: the
product trains and runs the model in its own HIP kernels, sdpa_amd/csrc/
mlp.hip; this file is the checker they are compared with).

Model: y = W3 relu(W2 relu(W1 x + b1) + b2) + b3, parameters packed as
(W1 row-major, b1, W2 row-major, b2, W3, b3); loss (1/n) sum (y - t)^2;
Adam with beta1 0.9, beta2 0.999, eps 1e-8 and bias correction.  No
counterpart in the reference (it feeds the initial-point hook of
Solutions::initialize, sdpa_dataset.cpp:47-74).  Pinned by central finite
differences of the loss (tests/test_mlp_oracle.py).
"""
import numpy as np


def unpack(theta, nin, nh):
    o = 0
    W1 = theta[o:o + nh * nin].reshape(nh, nin); o += nh * nin
    b1 = theta[o:o + nh]; o += nh
    W2 = theta[o:o + nh * nh].reshape(nh, nh); o += nh * nh
    b2 = theta[o:o + nh]; o += nh
    W3 = theta[o:o + nh]; o += nh
    b3 = theta[o]
    return W1, b1, W2, b2, W3, b3


def nparams(nin, nh):
    return nh * nin + nh + nh * nh + nh + nh + 1


def forward(theta, X, nin, nh):
    W1, b1, W2, b2, W3, b3 = unpack(theta, nin, nh)
    Z1 = X @ W1.T + b1
    Z2 = np.maximum(Z1, 0) @ W2.T + b2
    return np.maximum(Z2, 0) @ W3 + b3, Z1, Z2


def loss_grad(theta, X, t, nin, nh):
    W1, b1, W2, b2, W3, b3 = unpack(theta, nin, nh)
    n = X.shape[0]
    y, Z1, Z2 = forward(theta, X, nin, nh)
    e = y - t
    loss = float(e @ e) / n
    dy = 2.0 * e / n
    H2 = np.maximum(Z2, 0)
    gW3 = dy @ H2
    gb3 = dy.sum()
    dZ2 = np.outer(dy, W3) * (Z2 > 0)
    H1 = np.maximum(Z1, 0)
    gW2 = dZ2.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (Z1 > 0)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g


def adam_step(theta, g, m1, m2, t, lr):
    """One Adam step at step number t (1-based); returns new (theta, m1, m2)."""
    m1 = 0.9 * m1 + 0.1 * g
    m2 = 0.999 * m2 + 0.001 * g * g
    c1 = 1.0 - 0.9 ** t
    c2 = 1.0 - 0.999 ** t
    return theta - lr * (m1 / c1) / (np.sqrt(m2 / c2) + 1e-8), m1, m2
