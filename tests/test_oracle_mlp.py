"""Pin the MLP oracle's gradient by central finite differences (CPU, FP64)."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = 40
    p = rng.standard_normal(8641) * 0.2
    x = rng.standard_normal((n, 3))
    t = rng.standard_normal(n)
    w = rng.uniform(0.1, 1.0, n)
    loss, g = mlp.loss_and_grad(p, x, t, w)
    for idx in rng.choice(8641, 25, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (mlp.loss_and_grad(pp, x, t, w)[0] - mlp.loss_and_grad(pm, x, t, w)[0]) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])
