"""GPU parity: block Krylov solver and dense FP64 linear algebra against direct solves / the oracle.

Mirrors the reference's tests/test_krylov_block.rs:7-97 on the closed-form system of
prototype/02-7-krylov_testing_setup.py (n=40, nrhs=12), tolerances as there.
"""
import numpy as np
import pytest
import torch

from oracle import krylov as okr

pytestmark = pytest.mark.gpu


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def system(dev, n=40, nrhs=12):
    A, b = okr.closed_form_test_system(n, nrhs)
    return A, b, T(A, dev), T(b, dev)


def test_krylov_small_deterministic(dev):
    import hesship
    A, b, dA, db = system(dev)
    calls = []
    st = {}
    x = hesship.krylov_block(lambda v: (calls.append(1), dA @ v)[1], db, None, 1e-10, 30, 6, 1e-13, stats=st)
    x = x.cpu().numpy()
    x_ref = np.linalg.solve(np.eye(40) + A, b)
    np.testing.assert_allclose(x, x_ref, rtol=1e-6, atol=1e-8)
    assert np.abs(x + A @ x - b).max() < 1e-6
    # same trajectory length as the oracle restatement
    so = {}
    okr.krylov_block(lambda v: A @ v, b, None, 1e-10, 30, 6, 1e-13, stats=so)
    assert st["cycles"] == so["cycles"] and st["restarts"] == so["restarts"] == 0


def test_krylov_with_restart(dev):
    import hesship
    A, b, dA, db = system(dev)
    st = {}
    x = hesship.krylov_block(lambda v: dA @ v, db, None, 1e-10, 60, 3, 1e-13, stats=st).cpu().numpy()
    x_ref = np.linalg.solve(np.eye(40) + A, b)
    np.testing.assert_allclose(x, x_ref, rtol=1e-5, atol=1e-6)
    assert np.abs(x + A @ x - b).max() < 1e-5
    assert st["restarts"] >= 1


def test_krylov_nonconvergence_is_an_error(dev):
    import hesship
    A, b, dA, db = system(dev)
    with pytest.raises(hesship.KrylovNotConverged):
        hesship.krylov_block(lambda v: dA @ v, db, None, 1e-10, 2, 14, 1e-13)


def test_krylov_initial_guess_and_zero_rhs(dev):
    import hesship
    A, b, dA, db = system(dev, 20, 3)
    x_ref = np.linalg.solve(np.eye(20) + A, b)
    x = hesship.krylov_block(lambda v: dA @ v, db, T(x_ref, dev), 1e-10, 30, 6, 1e-13).cpu().numpy()
    np.testing.assert_allclose(x, x_ref, rtol=1e-8, atol=1e-10)
    z = hesship.krylov_block(lambda v: dA @ v, torch.zeros_like(db), None, 1e-10, 30, 6, 1e-13)
    assert torch.count_nonzero(z).item() == 0


def test_krylov_larger_block_matches_oracle(dev):
    """n = 300, nset = 20 (several orthogonalisation chunks), random non-symmetric A."""
    import hesship
    rng = np.random.default_rng(0)
    n, ns = 300, 20
    A = 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal((n, ns))
    dA = T(A, dev)
    st = {}
    x = hesship.krylov_block(lambda v: dA @ v, T(b, dev), None, 1e-9, 60, 14, 1e-14, stats=st).cpu().numpy()
    so = {}
    xo = okr.krylov_block(lambda v: A @ v, b, None, 1e-9, 60, 14, 1e-14, stats=so)
    # convergence is on the trial-vector norm (tol 1e-9), so the solution error is ~1e-8
    np.testing.assert_allclose(x, np.linalg.solve(np.eye(n) + A, b), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(x, xo, rtol=1e-6, atol=1e-8)
    assert st["cycles"] == so["cycles"]


@pytest.mark.parametrize("n,m", [(50, 1), (200, 37), (500, 300)])
def test_orth_block_matches_oracle(dev, n, m):
    """Columns over six orders of magnitude, exact and near (1e-9) dependencies, and m = 300
    (three Gram groups): same kept set, vectors and norms as the serial MGS of the oracle."""
    import hesship
    rng = np.random.default_rng(m)
    V = rng.standard_normal((n, m)) * np.exp(rng.uniform(-7, 0, m))
    if m > 10:
        V[:, 3] = 0.5 * V[:, 0] + V[:, 1]
        V[:, 7] = V[:, 2] + 1e-9 * rng.standard_normal(n)
        V[:, m - 1] = V[:, 5] - 2.0 * V[:, m - 2]
    R, nr = hesship.orth_block(T(V, dev), 1e-14)
    Ro, no = okr.orth_block(V, 1e-14)
    assert R.shape[1] == Ro.shape[1] == len(nr)
    np.testing.assert_allclose(R.cpu().numpy(), Ro, rtol=0, atol=1e-12 * np.abs(Ro).max())
    np.testing.assert_allclose(nr, no, rtol=1e-12)
    Rn = R.cpu().numpy() / np.sqrt(nr)
    assert np.abs(Rn.T @ Rn - np.eye(len(nr))).max() < 1e-12


@pytest.mark.parametrize("nd,nrhs", [(1, 1), (63, 5), (64, 64), (130, 7), (333, 90)])
def test_dense_solve(dev, nd, nrhs):
    import hesship
    rng = np.random.default_rng(nd)
    H = rng.standard_normal((nd, nd)) + nd * 0.1 * np.eye(nd)
    G = rng.standard_normal((nd, nrhs))
    X = hesship.solve(T(H, dev), T(G, dev)).cpu().numpy()
    np.testing.assert_allclose(H @ X, G, rtol=1e-9, atol=1e-9)
    if nd == 1:
        return
    # pivoting matters: a zero leading entry
    H2 = H.copy(); H2[0, 0] = 0.0
    X2 = hesship.solve(T(H2, dev), T(G, dev)).cpu().numpy()
    np.testing.assert_allclose(H2 @ X2, G, rtol=1e-8, atol=1e-8)


def test_dense_solve_singular_is_an_error(dev):
    import hesship
    H = np.ones((5, 5))
    with pytest.raises(hesship.HesshipError):
        hesship.solve(T(H, dev), T(np.ones((5, 1)), dev))


@pytest.mark.parametrize("nd", [3, 64, 150, 257])
def test_cholesky_and_trsm(dev, nd):
    import hesship
    rng = np.random.default_rng(nd)
    a = rng.standard_normal((nd, nd))
    S = a @ a.T + nd * np.eye(nd)
    L = hesship.cholesky(T(S, dev)).cpu().numpy()
    np.testing.assert_allclose(L, np.linalg.cholesky(S), rtol=1e-10, atol=1e-10)
    X = rng.standard_normal((nd, 11))
    np.testing.assert_allclose(hesship.trsm_lower(T(L, dev), T(X, dev)).cpu().numpy(),
                               np.linalg.solve(L, X), rtol=1e-9, atol=1e-9)
    Y = rng.standard_normal((37, nd))
    np.testing.assert_allclose(hesship.trsm_lower(T(L, dev), T(Y, dev), side_right=True).cpu().numpy(),
                               np.linalg.solve(L, Y.T).T, rtol=1e-9, atol=1e-9)
    with pytest.raises(hesship.HesshipError):
        hesship.cholesky(T(-S, dev))


@pytest.mark.parametrize("nd", [1, 2, 7, 64, 151, 400])
def test_eigh_jacobi(dev, nd):
    """The device Jacobi eigensolver against numpy.linalg.eigh: eigenvalues to 1e-13 of the
    spectral radius, orthonormal eigenvectors reproducing H, only the lower triangle read; a
    spectrum with a degenerate pair and eigenvalues of both signs over 12 decades."""
    import hesship
    rng = np.random.default_rng(nd)
    Q = np.linalg.qr(rng.standard_normal((nd, nd)))[0]
    lam = np.sign(rng.standard_normal(nd)) * 10 ** rng.uniform(-10, 2, nd)
    if nd >= 4:
        lam[1] = lam[2]
    H = (Q * lam) @ Q.T
    H = 0.5 * (H + H.T)
    Hin = H.copy()
    Hin[np.triu_indices(nd, 1)] = 1e300  # the upper triangle is not read
    w, V = hesship.eigh(T(Hin, dev))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    ref = np.linalg.eigvalsh(H)
    sc = np.abs(ref).max()
    assert np.abs(w - ref).max() <= 1e-13 * sc
    assert np.abs(V.T @ V - np.eye(nd)).max() < 1e-13
    assert np.abs((V * w) @ V.T - H).max() <= 1e-13 * sc
    assert bool(np.all(np.diff(w) >= 0))
