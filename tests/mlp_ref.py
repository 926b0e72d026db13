"""fp64 restatement of the MLP cost model (test infrastructure)."""
import numpy as np


def unpack(p, nin, h1, h2):
    o = 0
    W1 = p[o:o + h1 * nin].reshape(h1, nin); o += h1 * nin
    b1 = p[o:o + h1]; o += h1
    W2 = p[o:o + h2 * h1].reshape(h2, h1); o += h2 * h1
    b2 = p[o:o + h2]; o += h2
    W3 = p[o:o + h2]; o += h2
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def fwd64(p, X, nin, h1, h2):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, h1, h2)
    z1 = X @ W1.T + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0)
    return a2 @ W3 + b3, (z1, a1, z2, a2)


def loss64(p, X, t, nin, h1, h2):
    y, _ = fwd64(p, X, nin, h1, h2)
    return 0.5 * np.sum((y - t) ** 2) / len(t)


def grad64(p, X, t, nin, h1, h2):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, h1, h2)
    y, (z1, a1, z2, a2) = fwd64(p, X, nin, h1, h2)
    dy = (y - t) / len(t)
    dW3 = a2.T @ dy
    db3 = dy.sum()
    dz2 = np.outer(dy, W3) * (z2 > 0)
    dW2 = dz2.T @ a1
    db2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    dW1 = dz1.T @ X
    db1 = dz1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]])
