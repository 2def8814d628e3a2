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


def test_adam_matches_torch_optimizer():
    """The oracle's Adam equals torch.optim.Adam (an independent implementation) step by step."""
    import torch
    rng = np.random.default_rng(1)
    n = 30
    p0 = rng.standard_normal(8641) * 0.2
    x = rng.standard_normal((n, 3))
    t = rng.standard_normal(n)
    w = rng.uniform(0.1, 1.0, n)
    p_or, loss_or = mlp.adam_train(p0, x, t, w, epochs=4, lr=2e-3)
    pt = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=2e-3, betas=(0.9, 0.999), eps=1e-8)
    tt, wt = torch.as_tensor(t), torch.as_tensor(w)
    for _ in range(4):
        opt.zero_grad()
        y = mlp.forward(pt, x)
        loss = (wt * (y - tt) ** 2).sum() / wt.sum()
        loss.backward()
        opt.step()
    assert abs(loss.item() - loss_or) <= 1e-12 * abs(loss_or)
    assert np.abs(pt.detach().numpy() - p_or).max() <= 1e-12
