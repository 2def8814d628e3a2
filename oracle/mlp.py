"""Oracle: the learned-preconditioner MLP in FP64 (torch, CPU).  Test infrastructure only.

No reference counterpart (the reference has no learned component).  Restates the network of
csrc/mlp.hip with the same flat parameter layout: W1 [64, 3] column-major (W1[j + 64 f]), b1
[64], W2 [64, 64] (W2[k + 64 j]: out k, in j), b2, W3, b3, W4 [64], b4 [1]; ReLU on the three
hidden layers; loss = sum_n w_n (f(x_n) - t_n)^2 / sum_n w_n.
"""
import numpy as np
import torch

H, NIN = 64, 3


def unpack(p):
    p = torch.as_tensor(p, dtype=torch.float64)
    o = 0

    def take(n):
        nonlocal o
        v = p[o:o + n]
        o += n
        return v
    W1 = take(NIN * H).reshape(NIN, H).T
    b1 = take(H)
    W2 = take(H * H).reshape(H, H).T
    b2 = take(H)
    W3 = take(H * H).reshape(H, H).T
    b3 = take(H)
    W4 = take(H)
    b4 = take(1)
    assert o == p.numel()
    return W1, b1, W2, b2, W3, b3, W4, b4


def forward(p, x):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    h = torch.relu(torch.as_tensor(x, dtype=torch.float64) @ W1.T + b1)
    h = torch.relu(h @ W2.T + b2)
    h = torch.relu(h @ W3.T + b3)
    return h @ W4 + b4


def loss_and_grad(p, x, t, w):
    pp = torch.tensor(np.asarray(p, dtype=np.float64), requires_grad=True)
    y = forward(pp, x)
    t = torch.as_tensor(t, dtype=torch.float64)
    w = torch.as_tensor(w, dtype=torch.float64)
    loss = (w * (y - t) ** 2).sum() / w.sum()
    loss.backward()
    return loss.item(), pp.grad.numpy()


def adam_train(p, x, t, w, epochs, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """Full-batch Adam (Kingma & Ba 2015, bias-corrected moments) from zero moments, as the
    library's hh_mlp_train: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
    p -= lr (m / (1-b1^k)) / (sqrt(v / (1-b2^k)) + eps).  Returns (params, loss of the last step)."""
    p = np.asarray(p, dtype=np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    loss = 0.0
    for k in range(1, epochs + 1):
        loss, g = loss_and_grad(p, x, t, w)
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        p -= lr * (m / (1 - beta1 ** k)) / (np.sqrt(v / (1 - beta2 ** k)) + eps)
    return p, loss
