"""Pins of the TPSS meta-GGA in the functional oracle (oracle/xcfunc.py: MGGA_X_TPSS 202,
MGGA_C_TPSS 231, HYB_MGGA_XC_TPSS0 396) by what the TPSS construction and the literature fix:

* exchange: exact for the hydrogen atom (E_x = -5/16, which fixes c and e together with the
  uniform limit), the uniform-gas limit (alpha = 1, p = 0: Slater), Oliver-Perdew spin scaling;
* correlation: one-electron self-correlation free (E_c[H] = 0), the zero-gradient limit (PW92),
  the He atom at the Hartree-Fock density (LSD -0.1125, PBE -0.0420, TPSS -0.0431, Table I of
  Tao et al., Phys. Rev. Lett. 91, 146401 (2003), and Perdew et al., Phys. Rev. Lett. 77, 3865);
* PBE correlation of the H atom (-0.0060) and PW92 LSD (-0.0222).
The coefficients of C(zeta, xi) and d = 2.8 enter through the He value only ("parity partly
pinned" for them: the H and uniform pins are independent of them)."""
import math

import numpy as np
import pytest
import torch

from oracle import xcfunc as X


def _radial(N=4000, rmax=25.0):
    h = rmax / N
    r = np.arange(1, N + 1) * h
    return r, 4 * np.pi * r * r * h, h


def _hydrogen():
    r, w, _ = _radial()
    n = np.exp(-2 * r) / math.pi
    grad = -2 * n
    return r, w, n, grad


def _helium_hf():
    """Closed-shell HF 1s orbital of He on a radial grid (tridiagonal finite differences,
    self-consistent Hartree potential of one electron: F = h + J)."""
    from scipy.linalg import eigh_tridiagonal
    r, w, h = _radial(6000)
    phi = np.exp(-1.6875 * r)
    phi /= np.sqrt((phi ** 2 * w).sum())
    for _ in range(60):
        rho1 = phi ** 2
        q = np.cumsum(w * rho1)
        outer = np.cumsum((w * rho1 / r)[::-1])[::-1]
        V = -2 / r + q / r + outer - w * rho1 / r
        eps, u = eigh_tridiagonal(1 / h ** 2 + V, -0.5 / h ** 2 * np.ones(len(r) - 1), select="i",
                                  select_range=(0, 0))
        new = np.abs(u[:, 0]) / (np.sqrt(4 * np.pi * h) * r)
        phi = 0.5 * phi + 0.5 * new / np.sqrt((new ** 2 * w).sum())
    assert eps[0] == pytest.approx(-0.91796, abs=2e-4)   # the HF 1s eigenvalue of He
    n = 2 * phi ** 2
    return r, w, n, 2 * 2 * phi * np.gradient(phi, h)


def _t(x):
    return torch.tensor(np.asarray(x, dtype=np.float64))


def test_hydrogen_exchange_is_exact():
    r, w, n, grad = _hydrogen()
    s = grad ** 2
    tau = s / (8 * n)     # one orbital: tau = tau_W
    ex = X.tpss_x_spin(_t(n), _t(s), _t(tau)).numpy()
    assert (ex * w).sum() == pytest.approx(-0.3125, abs=2e-6)
    # through the raw polarised evaluation: rho_b = 0
    rho = np.zeros((len(r), 5, 2))
    rho[:, 0, 0], rho[:, 3, 0], rho[:, 4, 0] = n, grad, tau
    out = X.eval_raw(202, rho, 2, 0)
    assert (out[:, 0] * n * w).sum() == pytest.approx(-0.3125, abs=2e-6)


def test_hydrogen_correlation():
    r, w, n, grad = _hydrogen()
    rho = np.zeros((len(r), 5, 2))
    rho[:, 0, 0], rho[:, 3, 0], rho[:, 4, 0] = n, grad, grad ** 2 / (8 * n)
    ec = X.eval_raw(231, rho, 2, 0)[:, 0] * n
    assert np.abs(ec).max() < 1e-14                       # self-correlation free
    one = torch.ones(len(r), dtype=torch.float64)
    pbe = (n * X.pbe_c_eps(_t(n), one, _t(grad ** 2)).numpy() * w).sum()
    assert pbe == pytest.approx(-0.0060, abs=1e-4)
    rs = (3 / (4 * math.pi * n)) ** (1 / 3)
    lsd = (n * X.pw92_eps(_t(rs), one).numpy() * w).sum()
    assert lsd == pytest.approx(-0.0222, abs=1e-4)


def test_helium_correlation():
    r, w, n, grad = _helium_hf()
    one = torch.ones(len(r), dtype=torch.float64)
    rs = (3 / (4 * math.pi * n)) ** (1 / 3)
    assert (n * X.pw92_eps(_t(rs), 0 * one).numpy() * w).sum() == pytest.approx(-0.1125, abs=3e-4)
    assert (n * X.pbe_c_eps(_t(n), 0 * one, _t(grad ** 2)).numpy() * w).sum() == pytest.approx(-0.0420, abs=3e-4)
    rho = np.zeros((len(r), 5))
    rho[:, 0], rho[:, 3], rho[:, 4] = n, grad, grad ** 2 / (8 * n)
    ec = X.eval_raw(231, rho, 1, 0)[:, 0] * n
    assert (ec * w).sum() == pytest.approx(-0.0431, abs=4e-4)


def test_uniform_limits():
    n = np.array([1e-3, 0.02, 0.3, 2.0, 40.0])
    tau_unif = 0.3 * (3 * math.pi ** 2) ** (2 / 3) * n ** (5 / 3)
    rho = np.zeros((len(n), 5))
    rho[:, 0], rho[:, 4] = n, tau_unif
    slater = X.eval_raw(1, rho[:, :1], 1, 0)[:, 0]
    np.testing.assert_allclose(X.eval_raw(202, rho, 1, 0)[:, 0], slater, rtol=1e-13)
    rs = (3 / (4 * math.pi * n)) ** (1 / 3)
    pw = X.pw92_eps(_t(rs), _t(0 * n)).numpy()
    # at zero gradient tau_W = 0, so z = 0 and the correlation is PW92 for any tau
    for t in (tau_unif, 3 * tau_unif):
        rho[:, 4] = t
        np.testing.assert_allclose(X.eval_raw(231, rho, 1, 0)[:, 0], pw, rtol=1e-13)
    # PW92 against the paramagnetic/ferromagnetic fits at zeta = 0, 1
    np.testing.assert_allclose(pw, X.pw92_g(_t(rs), *X.PWMOD[0]).numpy(), rtol=1e-15)
    np.testing.assert_allclose(X.pw92_eps(_t(rs), _t(n * 0 + 1)).numpy(), X.pw92_g(_t(rs), *X.PWMOD[1]).numpy(),
                               rtol=1e-13)


def test_tpss_exchange_spin_scaling_and_tpss0():
    rng = np.random.default_rng(3)
    ng = 20
    rho = np.zeros((ng, 5, 2))
    rho[:, 0] = rng.uniform(0.05, 2.0, (ng, 2))
    rho[:, 1:4] = rng.normal(size=(ng, 3, 2)) * rho[:, :1]
    rho[:, 4] = (rho[:, 1:4] ** 2).sum(1) / (8 * rho[:, 0]) * rng.uniform(1.1, 3.0, (ng, 2))
    ex = X.eval_raw(202, rho, 2, 0)[:, 0] * rho[:, 0].sum(1)
    parts = sum(0.5 * X.eval_raw(202, 2 * rho[..., s], 1, 0)[:, 0] * 2 * rho[:, 0, s] for s in range(2))
    np.testing.assert_allclose(ex, parts, rtol=1e-12)
    tp0 = X.eval_raw(396, rho, 2, 0)[:, 0]
    np.testing.assert_allclose(tp0, 0.75 * X.eval_raw(202, rho, 2, 0)[:, 0] + X.eval_raw(231, rho, 2, 0)[:, 0],
                               rtol=1e-13)
    assert X.EXX[396] == 0.25
