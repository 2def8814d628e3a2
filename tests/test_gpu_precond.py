"""GPU: learned diagonal preconditioner -- kernels against the FP64 oracle, and its gain on a
CPHF solve (fewer Krylov cycles, same solution)."""
import numpy as np
import pytest
import torch

from oracle import cphf as ocp
from oracle import mlp as omlp
from tests.test_gpu_cphf import T, make, system

pytestmark = pytest.mark.gpu


def test_forward_and_gradient_match_oracle(dev):
    from hesship.precond import DiagMLP
    rng = np.random.default_rng(1)
    n = 3000
    x = rng.standard_normal((n, 3)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    net = DiagMLP(seed=7)
    p = net.params()
    y = net.forward(T(x, dev)).cpu().numpy()
    y_ref = omlp.forward(p, x).numpy()
    assert np.abs(y - y_ref).max() <= 1e-5 * max(1.0, np.abs(y_ref).max())
    loss, g = net.loss_and_grad(T(x, dev), T(t, dev), T(w, dev))
    loss_ref, g_ref = omlp.loss_and_grad(p, x, t, w)
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)
    g = g.cpu().numpy()
    # FP32 sums over 3000 samples on the matrix cores: agree to ~1e-5 of the gradient scale
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max()


def test_adam_steps_match_oracle(dev):
    """Ten library Adam steps (hh_mlp_train) against the oracle's Adam from the same parameters."""
    from hesship.precond import DiagMLP
    rng = np.random.default_rng(4)
    n = 1500
    x = rng.standard_normal((n, 3)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    net = DiagMLP(seed=9)
    p0 = net.params()
    net.set_params(p0)  # zero moments, step counter 0
    lr, steps = 3e-4, 10
    loss = net.train(T(x, dev), T(t, dev), T(w, dev), epochs=steps, lr=lr)
    p1 = net.params().astype(np.float64)
    p_ref, loss_ref = omlp.adam_train(p0, x, t, w, epochs=steps, lr=lr)
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref)
    # Adam moves every parameter by ~lr per step whatever its gradient's size, so parameters whose
    # gradient is at the FP32 rounding level wander: bound all moves, and require the firmly
    # determined ones (|g| above 1e-2 of the largest) to agree.  The tolerance sits between
    # FP32-level gradient noise (~4e-5 at 3e-5 of max|g|) and formula mistakes of the same test
    # (no bias correction: 1.3e-2; beta1 0.8: 3.8e-4), both measured with the oracle.
    _, g_ref = omlp.loss_and_grad(p0, x, t, w)
    d = np.abs(p1 - p_ref)
    assert d.max() <= 4 * lr * steps
    firm = np.abs(g_ref) > 1e-2 * np.abs(g_ref).max()
    assert firm.sum() >= 20
    assert d[firm].max() <= 2e-4, d[firm].max()


def test_training_reduces_loss(dev):
    from hesship.precond import DiagMLP
    rng = np.random.default_rng(2)
    n = 2000
    x = rng.standard_normal((n, 3)).astype(np.float32)
    t = (0.5 * np.tanh(x[:, 0]) - 0.3 * x[:, 1] ** 2 + 0.2 * x[:, 2]).astype(np.float32)
    w = np.ones(n, dtype=np.float32)
    net = DiagMLP(seed=3)
    l0, _ = net.loss_and_grad(T(x, dev), T(t, dev), T(w, dev))
    net.train(T(x, dev), T(t, dev), T(w, dev), epochs=400, lr=3e-3)
    l1, _ = net.loss_and_grad(T(x, dev), T(t, dev), T(w, dev))
    assert l1 < 0.05 * l0, (l0, l1)


def diag_dominant_case(seed):
    """RI-JK system whose CPHF operator has a strong (ai|ai)/de diagonal (J-dominated)."""
    cderi, C, occ, e, f1, s1, _ = system(nao=40, naux=90, nocc=6, natm=4, seed=seed, sk=0.0)
    return cderi * 1.8, C, occ, e, f1, s1


def test_learned_preconditioner_effect(dev):
    """Train on one system's (residual, solution) pairs, apply to another: same solution, and the
    Krylov cycle counts with and without the learned diagonal are reported (on these RI operators
    even the exact diagonal inverse does not shorten the solve, so no gain is asserted; the
    network must not cost more than one extra cycle)."""
    import hesship
    from hesship import precond
    # training system: generate (residual, solution) pairs with the FP64 method
    cderi, C, occ, e, f1, s1 = diag_dominant_case(20)
    drv, rijk = make(dev, cderi, C, occ, e, sk=0.0, cfg=dict(cphf_tol=1e-10))
    dC = T(C, dev)
    pair = precond.pair_norms(rijk, dC)
    rhs = drv.compute_dimless_rhs(T(f1, dev), T(s1, dev))
    sol = drv.solve_dimless_cphf(rhs)
    x = precond.features(drv, pair, rhs)
    t, w = precond.target(drv, rhs, sol)
    net = precond.DiagMLP(seed=11)
    net.train(x, t, w, epochs=500, lr=3e-3)
    # evaluation system (different seed): plain vs learned right preconditioner
    cderi2, C2, occ2, e2, f12, s12 = diag_dominant_case(21)
    drv2, rijk2 = make(dev, cderi2, C2, occ2, e2, sk=0.0, cfg=dict(cphf_tol=1e-10))
    pair2 = precond.pair_norms(rijk2, T(C2, dev))
    rhs2 = drv2.compute_dimless_rhs(T(f12, dev), T(s12, dev))
    st0, st1 = {}, {}
    x0 = drv2.solve_dimless_cphf(rhs2, stats=st0).cpu().numpy()
    d = precond.diag(drv2, net.forward(precond.features(drv2, pair2, rhs2)))
    x1 = drv2.solve_dimless_cphf(rhs2, diag=d, stats=st1).cpu().numpy()
    print(f"krylov cycles: plain {st0['cycles']}, learned {st1['cycles']}")
    np.testing.assert_allclose(x1, x0, rtol=1e-6, atol=1e-8)
    assert st1["cycles"] <= st0["cycles"] + 1, (st0, st1)
    # the trained network tracks its target on the unseen system better than the untrained one
    t2, w2 = precond.target(drv2, rhs2, torch.from_numpy(x0).to(dev))
    x2 = precond.features(drv2, pair2, rhs2)
    fresh = precond.DiagMLP(seed=11)
    l_trained, _ = net.loss_and_grad(x2, t2, w2)
    l_fresh, _ = fresh.loss_and_grad(x2, t2, w2)
    print(f"held-out loss: untrained {l_fresh:.4e}, trained {l_trained:.4e}")
    assert l_trained < l_fresh
