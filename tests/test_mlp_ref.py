"""The fp64 MLP restatement is pinned by central finite differences (CPU)."""
import numpy as np

from mlp_ref import grad64, loss64


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nin, h1, h2, B = 5, 7, 6, 9
    P = h1 * nin + h1 + h2 * h1 + h2 + h2 + 1
    p = rng.normal(0, 0.5, P)
    X = rng.normal(0, 1, (B, nin))
    t = rng.normal(0, 1, B)
    g = grad64(p, X, t, nin, h1, h2)
    eps = 1e-6
    for i in range(P):
        e = np.zeros(P); e[i] = eps
        fd = (loss64(p + e, X, t, nin, h1, h2) - loss64(p - e, X, t, nin, h1, h2)) / (2 * eps)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), i
