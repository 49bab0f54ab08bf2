"""TEST INFRASTRUCTURE ONLY: float64 numpy restatement of the cost-model MLP.

The HIP kernels in hacapk_hmvm_amd/csrc/hmvm_mlp.hip compute the forward pass,
the loss 0.5/n * sum (f(x_i) - y_i)^2 and its gradient for the network
nin -> nh -> nh -> 1 with ReLU activations on standardised inputs.  This module
computes the same quantities in float64 with plain matrix algebra so the tests
can check the GPU gradients.  It has no reference counterpart (the reference
has no learned component); it is pinned in tests/test_cost_model.py against
central finite differences of its own loss, which fail for a dropped term, a
wrong sign or a transposed weight.

Parameter vector layout (as in the HIP code): W1 (nh x nin, row-major), b1
(nh), W2 (nh x nh), b2 (nh), w3 (nh), b3 (1).
"""
import numpy as np


def unpack(params, nin, nh):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + nh * nin].reshape(nh, nin); o += nh * nin
    b1 = p[o:o + nh]; o += nh
    W2 = p[o:o + nh * nh].reshape(nh, nh); o += nh * nh
    b2 = p[o:o + nh]; o += nh
    w3 = p[o:o + nh]; o += nh
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, w3, b3


def forward(params, mean, stdev, X, nin, nh):
    W1, b1, W2, b2, w3, b3 = unpack(params, nin, nh)
    sd = np.where(np.asarray(stdev, dtype=np.float64) > 0, stdev, 1.0)
    Xs = (np.asarray(X, dtype=np.float64) - np.asarray(mean, dtype=np.float64)) / sd
    Z1 = Xs @ W1.T + b1
    H1 = np.maximum(Z1, 0.0)
    Z2 = H1 @ W2.T + b2
    H2 = np.maximum(Z2, 0.0)
    out = H2 @ w3 + b3
    return out, (Xs, H1, H2)


def loss_grad(params, mean, stdev, X, y, nin, nh):
    """(loss, gradient) of 0.5/n * sum (f(X_i) - y_i)^2."""
    W1, b1, W2, b2, w3, b3 = unpack(params, nin, nh)
    out, (Xs, H1, H2) = forward(params, mean, stdev, X, nin, nh)
    n = X.shape[0]
    e = (out - np.asarray(y, dtype=np.float64)) / n
    loss = 0.5 * np.sum((out - y) ** 2) / n
    gw3 = H2.T @ e
    gb3 = e.sum()
    D2 = np.outer(e, w3) * (H2 > 0)
    gW2 = D2.T @ H1
    gb2 = D2.sum(axis=0)
    D1 = (D2 @ W2) * (H1 > 0)
    gW1 = D1.T @ Xs
    gb1 = D1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g
