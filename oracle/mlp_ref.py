"""FP64 restatement of the plan selector's network -- TEST INFRASTRUCTURE ONLY.

The network of cuspfft_amd/csrc/mlp.hip: D -> H -> H -> 1 with ReLU, parameters
packed as W1 [H x D], b1 [H], W2 [H x H], b2 [H], W3 [H], b3 [1], loss
L = 1/(2B) sum_b (yhat_b - y_b)^2.  Used to check the HIP forward and backward
kernels; pinned itself by central finite differences (tests/test_selector.py).
"""
import numpy as np


def unpack(p, D, H):
    p = np.asarray(p, np.float64)
    i = 0
    W1 = p[i:i + H * D].reshape(H, D); i += H * D
    b1 = p[i:i + H]; i += H
    W2 = p[i:i + H * H].reshape(H, H); i += H * H
    b2 = p[i:i + H]; i += H
    W3 = p[i:i + H]; i += H
    b3 = p[i:i + 1]; i += 1
    assert i == len(p)
    return W1, b1, W2, b2, W3, b3


def forward(p, X, D, H):
    W1, b1, W2, b2, W3, b3 = unpack(p, D, H)
    X = np.asarray(X, np.float64)
    Z1 = X @ W1.T + b1
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ W2.T + b2
    A2 = np.maximum(Z2, 0)
    yhat = A2 @ W3 + b3[0]
    return yhat, (X, Z1, A1, Z2, A2)


def loss_grad(p, X, y, D, H):
    W1, b1, W2, b2, W3, b3 = unpack(p, D, H)
    yhat, (X, Z1, A1, Z2, A2) = forward(p, X, D, H)
    y = np.asarray(y, np.float64)
    B = len(y)
    d = yhat - y
    loss = float(d @ d / (2 * B))
    dy = d / B
    gW3 = A2.T @ dy
    gb3 = np.array([dy.sum()])
    dZ2 = np.outer(dy, W3) * (Z2 > 0)
    gW2 = dZ2.T @ A1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (Z1 > 0)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
