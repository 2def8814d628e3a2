"""Pin the Krylov oracle against direct solves (reference tests/test_krylov_block.rs:7-97)."""
import numpy as np
import pytest

from oracle import krylov


def test_closed_form_direct_solve():
    A, b = krylov.closed_form_test_system(40, 12)
    x_ref = np.linalg.solve(np.eye(40) + A, b)
    st = {}
    x = krylov.krylov_block(lambda v: A @ v, b, None, 1e-10, 30, 6, 1e-13, stats=st)
    np.testing.assert_allclose(x, x_ref, rtol=1e-6, atol=1e-8)
    assert np.abs(x + A @ x - b).max() < 1e-6
    assert st["restarts"] == 0


def test_forced_restarts():
    A, b = krylov.closed_form_test_system(40, 12)
    x_ref = np.linalg.solve(np.eye(40) + A, b)
    st = {}
    x = krylov.krylov_block(lambda v: A @ v, b, None, 1e-10, 60, 3, 1e-13, stats=st)
    np.testing.assert_allclose(x, x_ref, rtol=1e-5, atol=1e-6)
    assert np.abs(x + A @ x - b).max() < 1e-5
    assert st["restarts"] >= 1


def test_nonconvergence_raises():
    A, b = krylov.closed_form_test_system(40, 12)
    with pytest.raises(krylov.KrylovNotConverged):
        krylov.krylov_block(lambda v: A @ v, b, None, 1e-10, 2, 14, 1e-13)


def test_initial_guess_and_zero_rhs():
    A, b = krylov.closed_form_test_system(20, 3)
    x_ref = np.linalg.solve(np.eye(20) + A, b)
    x = krylov.krylov_block(lambda v: A @ v, b, x_ref.copy(), 1e-10, 30, 6, 1e-13)
    np.testing.assert_allclose(x, x_ref, rtol=1e-8, atol=1e-10)
    x0 = krylov.krylov_block(lambda v: A @ v, np.zeros_like(b), None, 1e-10, 30, 6, 1e-13)
    np.testing.assert_array_equal(x0, 0.0)


def test_orth_block_drops_dependent():
    rng = np.random.default_rng(0)
    v = rng.standard_normal((10, 3))
    v = np.concatenate([v, v[:, :1] + v[:, 1:2]], axis=1)
    out, nrm = krylov.orth_block(v, 1e-14)
    assert out.shape[1] == 3
    g = out.T @ out
    np.testing.assert_allclose(g - np.diag(np.diag(g)), 0.0, atol=1e-12)
    np.testing.assert_allclose(np.diag(g), nrm, rtol=1e-12)
