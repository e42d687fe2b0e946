"""CPU restatement of the rank model's network (test infrastructure): the
forward pass, the mean squared error and its gradient by back-propagation,
in the parameter order of flam_amd/csrc/mlp_kernels.hip:

    W1 (H x F, row-major), b1 (H), W2 (H x H), b2 (H), w3 (H), b3 (1),
    y = w3 . tanh(W2 tanh(W1 x + b1) + b2) + b3,  loss = mean (y - t)^2.

Pinned by central finite differences of the loss (tests/test_oracle.py)."""
import numpy as np

F, H = 8, 32
NPARAMS = H * F + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=float)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    A1 = np.tanh(X @ W1.T + b1)
    A2 = np.tanh(A1 @ W2.T + b2)
    return A2 @ w3 + b3, A1, A2


def loss_grad(p, X, t):
    """(loss, gradient) of mean (y - t)^2."""
    X = np.asarray(X, dtype=float)
    n = X.shape[0]
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, A1, A2 = forward(p, X)
    e = y - t
    loss = float(np.mean(e * e))
    d3 = 2.0 * e / n                                  # dloss/dy
    D2 = d3[:, None] * w3[None, :] * (1.0 - A2 * A2)  # back through tanh
    D1 = (D2 @ W2) * (1.0 - A1 * A1)
    g = np.concatenate([(D1.T @ X).ravel(), D1.sum(0), (D2.T @ A1).ravel(), D2.sum(0),
                        A2.T @ d3, [d3.sum()]])
    return loss, g
