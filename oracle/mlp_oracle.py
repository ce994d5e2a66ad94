"""FP64 numpy restatement of the SPR-scoring MLP (TEST INFRASTRUCTURE ONLY:
the checker for the GPU kernels of phylocaml_amd/csrc/mlp.hip).

Layers: Z1 = X W1 + b1, H1 = gelu(Z1), Z2 = H1 W2 + b2, H2 = gelu(Z2),
y = H2 w3 + b3, loss = 0.5 mean((y - t)^2); gelu(x) = x Phi(x) (the erf form).
Pinned by central finite differences of the loss (tests/test_oracle_mlp.py).
"""
import numpy as np
from scipy.special import erf


def unpack(p, n_in, H):
    o = 0
    W1 = p[o:o + n_in * H].reshape(n_in, H); o += n_in * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def forward(p, X, n_in, H):
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64), n_in, H)
    Z1 = X @ W1 + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = gelu(Z2)
    return H2 @ w3 + b3, (Z1, H1, Z2, H2)


def loss_and_grad(p, X, t, n_in, H):
    X = np.asarray(X, np.float64)
    t = np.asarray(t, np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64), n_in, H)
    y, (Z1, H1, Z2, H2) = forward(p, X, n_in, H)
    B = X.shape[0]
    d = y - t
    loss = 0.5 * np.mean(d * d)
    dy = d / B
    gw3 = H2.T @ dy
    gb3 = dy.sum()
    dZ2 = np.outer(dy, w3) * gelu_grad(Z2)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * gelu_grad(Z1)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g
