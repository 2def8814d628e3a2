"""Pins of the functional oracle (oracle/xcfunc.py) by what the mathematics and the literature fix.

LibXC is not available here; each functional is checked against closed forms, limits that
fix its constants, an independent parametrisation of the same Monte Carlo data, spin-scaling
identities, and finite differences of its own energy (which also fixes LibXC's column order).
"""
import math

import numpy as np
import pytest
import torch

from oracle import xceff
from oracle import xcfunc as X

IDS = sorted(X.FAMILY)


def _rho_unpol(rng, ng, fam=1):
    """[ng, 4] (rho, grad) or, for a meta-GGA, [ng, 5] with tau above the von Weizsaecker bound
    |grad rho|^2 / 8 rho (the smooth branch of the functional)."""
    r = np.empty((ng, 5 if fam == 2 else 4))
    r[:, 0] = rng.uniform(0.05, 3.0, ng)
    r[:, 1:4] = rng.normal(size=(ng, 3)) * r[:, :1] * rng.uniform(0.1, 1.5, (ng, 1))
    if fam == 2:
        tw = (r[:, 1:4] ** 2).sum(1) / (8 * r[:, 0])
        r[:, 4] = tw * rng.uniform(1.2, 4.0, ng) + 0.05 * r[:, 0] ** (5 / 3)
    return r


def _rho_pol(rng, ng, fam=1):
    return np.stack([_rho_unpol(rng, ng, fam), _rho_unpol(rng, ng, fam)], -1)


def _t(x):
    return torch.tensor(np.asarray(x, dtype=np.float64))


def _pw92(rs, A, a1, b1, b2, b3, b4):
    """Perdew-Wang 1992 G(rs) (Phys. Rev. B 45, 13244, eq. 10)."""
    q = 2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** 2)
    return -2 * A * (1 + a1 * rs) * np.log1p(1 / q)


PW92_PARA = (0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294)
PW92_FERRO = (0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517)
PW92_MALPHA = (0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671)   # -alpha_c


def test_slater_closed_form():
    rho = np.linspace(0.01, 5.0, 40)[:, None]
    out = X.eval_raw(1, rho, 1, 2)
    cx = 0.75 * (3.0 / math.pi) ** (1.0 / 3.0)   # 0.7386 (Dirac)
    zk = -cx * rho[:, 0] ** (1.0 / 3.0)
    np.testing.assert_allclose(out[:, 0], zk, rtol=1e-13)
    np.testing.assert_allclose(out[:, 1], 4.0 / 3.0 * zk, rtol=1e-13)
    np.testing.assert_allclose(out[:, 2], 4.0 / 9.0 * zk / rho[:, 0], rtol=1e-12)
    assert abs(cx - 0.7385587663820224) < 1e-15


@pytest.mark.parametrize("xc_id", [1, 106])
def test_exchange_spin_scaling(xc_id):
    """E_x[rho_a, rho_b] = (E_x[2 rho_a] + E_x[2 rho_b]) / 2 (Oliver-Perdew)."""
    rng = np.random.default_rng(11)
    rp = _rho_pol(rng, 30)
    e_pol = X.eval_raw(xc_id, rp, 2, 0)[:, 0] * rp[:, 0].sum(-1)
    e_s = [X.eval_raw(xc_id, 2.0 * rp[:, :, s], 1, 0)[:, 0] * 2.0 * rp[:, 0, s] for s in range(2)]
    np.testing.assert_allclose(e_pol, 0.5 * (e_s[0] + e_s[1]), rtol=1e-13)


def test_b88_limits():
    r = _t([0.3, 1.0, 2.5])
    # no gradient: Slater
    np.testing.assert_allclose(X.b88_spin(r, torch.zeros_like(r) + 1e-300).numpy(), X.slater_spin(r).numpy(),
                               rtol=1e-15)
    # gradient expansion: (e - e_LDA) / (rho^(4/3) x^2) -> -beta
    x = 1e-3
    s = (x * r ** (4.0 / 3.0)) ** 2
    d = (X.b88_spin(r, s) - X.slater_spin(r)) / (r ** (4.0 / 3.0) * x * x)
    np.testing.assert_allclose(d.numpy(), -0.0042, rtol=1e-6)
    # large gradient: (e - e_LDA) / rho^(4/3) -> -x / (6 asinh x), independent of beta
    x = 1e8
    s = (x * r ** (4.0 / 3.0)) ** 2
    d = (X.b88_spin(r, s) - X.slater_spin(r)) / r ** (4.0 / 3.0)
    np.testing.assert_allclose(d.numpy(), -x / (6.0 * math.asinh(x)), rtol=1e-6)


def test_vwn_high_density_limit():
    """eps_c -> A ln rs + C: A = (1 - ln 2)/pi^2 paramagnetic (half that ferromagnetic); C the
    exact second-order constant -0.046644 (fit V) and the RPA constant -0.0708 (RPA set);
    ferromagnetic exact -0.025599."""
    rs = _t([1e-10])
    a_para = (1.0 - math.log(2.0)) / math.pi ** 2
    assert abs(X.VWN5[0][0] - a_para) < 1e-7 and abs(X.VWN5[1][0] - a_para / 2) < 1e-7
    c = [X.vwn_curve(p, rs).item() - p[0] * math.log(1e-10) for p in X.VWN5]
    assert abs(c[0] + 0.046644) < 2e-5
    assert abs(c[1] + 0.025599) < 1e-4
    c_rpa = X.vwn_curve(X.VWN_RPA[0], rs).item() - X.VWN_RPA[0][0] * math.log(1e-10)
    assert abs(c_rpa + 0.0708) < 1e-4


def test_vwn5_against_pw92():
    """Fit V and Perdew-Wang 1992 parametrise the same Ceperley-Alder data."""
    rs = np.array([0.1, 0.5, 1.0, 2.0, 5.0, 10.0, 20.0, 50.0])
    t = _t(rs)
    np.testing.assert_allclose(X.vwn_curve(X.VWN5[0], t).numpy(), _pw92(rs, *PW92_PARA), atol=6e-4)
    np.testing.assert_allclose(X.vwn_curve(X.VWN5[1], t).numpy(), _pw92(rs, *PW92_FERRO), atol=2e-4)
    np.testing.assert_allclose(X.vwn_curve(X.VWN5[2], t).numpy(), -_pw92(rs, *PW92_MALPHA), atol=1.5e-3)


@pytest.mark.parametrize("params", ["VWN5", "VWN_RPA"])
def test_vwn_spin_interpolation(params):
    """eps(rs, 0) = eps_P, eps(rs, 1) = eps_F, d2 eps / d zeta2 at 0 = alpha_c (the third curve)."""
    P = getattr(X, params)
    rs = _t([0.7, 3.0, 12.0])
    eP, eF, eA = (X.vwn_curve(p, rs) for p in P)
    np.testing.assert_allclose(X.vwn_eps(P, rs, torch.zeros_like(rs)).numpy(), eP.numpy(), rtol=1e-14)
    np.testing.assert_allclose(X.vwn_eps(P, rs, torch.ones_like(rs)).numpy(), eF.numpy(), rtol=1e-14)
    h = 1e-3
    z = torch.full_like(rs, h)
    d2 = (X.vwn_eps(P, rs, z) - 2 * X.vwn_eps(P, rs, 0 * z) + X.vwn_eps(P, rs, -z)) / h ** 2
    np.testing.assert_allclose(d2.numpy(), eA.numpy(), rtol=1e-5)


def test_lyp_closed_shell_form():
    """Closed shell: -a rho/(1 + d rho^-1/3) - a b omega rho^2 [C_F rho^(8/3)
    + (5/12 - 7 delta/72) |grad rho|^2 - 11/24 |grad rho|^2] (Miehlich et al. 1989)."""
    rng = np.random.default_rng(3)
    r = _rho_unpol(rng, 25)
    rho, sig = r[:, 0], (r[:, 1:] ** 2).sum(1)
    a, b, c, d = X.LYP_A, X.LYP_B, X.LYP_C, X.LYP_D
    rm = rho ** (-1.0 / 3.0)
    omega = np.exp(-c * rm) / (1 + d * rm) * rho ** (-11.0 / 3.0)
    delta = c * rm + d * rm / (1 + d * rm)
    e = -a * rho / (1 + d * rm) - a * b * omega * rho ** 2 * (
        X.CF * rho ** (8.0 / 3.0) + (5.0 / 12 - 7.0 * delta / 72) * sig - 11.0 / 24 * sig)
    np.testing.assert_allclose(X.eval_raw(131, r, 1, 0)[:, 0] * rho, e, rtol=1e-13)


def test_lyp_vanishes_fully_polarised():
    """LYP correlation is zero for a one-spin density (e.g. the hydrogen atom)."""
    ra = _t([0.2, 1.0, 3.0])
    saa = _t([0.05, 0.7, 4.0])
    z = torch.zeros_like(ra)
    e = X.lyp(ra, z, saa, z, z)
    assert np.abs(e.numpy()).max() < 1e-16


@pytest.mark.parametrize("xc_id", IDS)
def test_unpolarised_is_polarised_at_equal_spins(xc_id):
    rng = np.random.default_rng(5)
    r = _rho_unpol(rng, 20, X.FAMILY[xc_id])
    rp = np.stack([r / 2, r / 2], -1)
    u = X.eval_raw(xc_id, r, 1, 1)
    p = X.eval_raw(xc_id, rp, 2, 1)
    np.testing.assert_allclose(p[:, 0], u[:, 0], rtol=1e-12)
    np.testing.assert_allclose(p[:, 1], u[:, 1], rtol=1e-11)   # vrho_a = vrho
    if X.GGA[xc_id]:
        # sigma = saa + 2 sab + sbb at equal spins: vsigma = (vs_aa + vs_ab + vs_bb) / 4
        np.testing.assert_allclose((p[:, 3] + p[:, 4] + p[:, 5]) / 4, u[:, 2], rtol=1e-10, atol=1e-14)
    if X.FAMILY[xc_id] == 2:
        # tau = tau_a + tau_b: vtau_a = vtau
        np.testing.assert_allclose(p[:, 6], u[:, 3], rtol=1e-10, atol=1e-14)


def test_b3lyp_composition():
    rng = np.random.default_rng(8)
    for nspin, r in ((1, _rho_unpol(rng, 15)), (2, _rho_pol(rng, 15))):
        parts = 0.08 * X.eval_raw(1, r, nspin, 0) + 0.72 * X.eval_raw(106, r, nspin, 0) \
            + 0.19 * X.eval_raw(8, r, nspin, 0) + 0.81 * X.eval_raw(131, r, nspin, 0)
        np.testing.assert_allclose(X.eval_raw(402, r, nspin, 0), parts, rtol=1e-13)


def _energy_of_components(xc_id, comp, nspin):
    """Energy density as a function of the density components (rho, grad rho[, tau]) of each spin."""
    fam = X.FAMILY[xc_id]
    if nspin == 1:
        v = [comp[:, 0]] + ([(comp[:, 1:4] ** 2).sum(1)] if fam >= 1 else []) + ([comp[:, 4]] if fam == 2 else [])
    else:
        v = [comp[:, 0, 0], comp[:, 0, 1]]
        if fam >= 1:
            ga, gb = comp[:, 1:4, 0], comp[:, 1:4, 1]
            v += [(ga * ga).sum(1), (ga * gb).sum(1), (gb * gb).sum(1)]
        if fam == 2:
            v += [comp[:, 4, 0], comp[:, 4, 1]]
    return X.energy_density(xc_id, *X._polarised_args(v, nspin, fam))


@pytest.mark.parametrize("nspin", [1, 2])
@pytest.mark.parametrize("xc_id", IDS)
def test_eff_against_direct_derivatives(xc_id, nspin):
    """vxc_eff / fxc_eff (LibXC columns -> XCFun order -> transform) equal the derivatives of
    the energy density in the density components themselves: pins the raw column order."""
    rng = np.random.default_rng(17 + xc_id + nspin)
    ng = 12
    fam = X.FAMILY[xc_id]
    rho = _rho_unpol(rng, ng, fam) if nspin == 1 else _rho_pol(rng, ng, fam)
    exc, vxc, fxc = X.eval_eff(xc_id, rho, nspin, "tau" if fam == 2 else "sigma")
    comp = torch.tensor(rho, requires_grad=True)
    e = _energy_of_components(xc_id, comp, nspin)
    (g,) = torch.autograd.grad(e.sum(), comp, create_graph=True)
    n = comp.numel() // ng
    gf = g.reshape(ng, n)
    H = torch.stack([torch.autograd.grad(gf[:, i].sum(), comp, retain_graph=True)[0].reshape(ng, n)
                     for i in range(n)], 1)
    # component order in vxc/fxc: (x, s) with x fastest for the reference's [ng, nvar, 2] shape
    if nspin == 1:
        np.testing.assert_allclose(vxc, g.detach().numpy(), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(fxc, H.detach().numpy(), rtol=1e-9, atol=1e-11)
    else:
        np.testing.assert_allclose(vxc, g.detach().numpy(), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(fxc.reshape(ng, n, n), H.detach().numpy(), rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(exc * (rho[:, 0] if nspin == 1 else rho[:, 0].sum(-1)), e.detach().numpy(),
                               rtol=1e-13)


@pytest.mark.parametrize("xc_id", IDS)
def test_raw_derivatives_against_finite_differences(xc_id):
    rng = np.random.default_rng(23)
    fam = X.FAMILY[xc_id]
    rho = _rho_pol(rng, 6, fam)
    v, _ = X._variables(rho, 2, fam)
    raw = X.eval_raw(xc_id, rho, 2, 2)
    nv = len(v)
    base = [t.clone() for t in v]

    def e_of(vals):
        return X.energy_density(xc_id, *X._polarised_args(vals, 2, fam)).numpy()

    h = 1e-5
    for i in range(nv):
        vp = [t.clone() for t in base]
        vm = [t.clone() for t in base]
        hi = h * torch.clamp(base[i].abs(), min=1e-2)
        vp[i] += hi
        vm[i] -= hi
        fd = (e_of(vp) - e_of(vm)) / (2 * hi.numpy())
        np.testing.assert_allclose(raw[:, 1 + i], fd, rtol=2e-6, atol=1e-8)


def test_thresholds():
    r = np.zeros((3, 4))
    r[:, 0] = [0.0, 5e-16, 1e-3]
    r[2, 1:] = 1e-4
    out = X.eval_raw(402, r, 1, 2)
    assert np.all(out[:2] == 0.0) and np.all(np.isfinite(out[2])) and out[2, 0] < 0


def test_transform_used_by_eff_is_the_pinned_one():
    """eval_eff delegates the chain rule to oracle.xceff (pinned against the reference's
    transform tests); an LDA functional fills the leading slice of SIGMA arrays."""
    rng = np.random.default_rng(2)
    r = _rho_unpol(rng, 7)
    exc, vxc, fxc = X.eval_eff(7, r, 1, "sigma")
    raw = X.eval_raw(7, r, 1, 2)
    np.testing.assert_array_equal(vxc[:, 0], raw[:, 1])
    assert np.all(vxc[:, 1:] == 0) and np.all(fxc[:, 1:, :] == 0) and np.all(fxc[:, :, 1:] == 0)
    exc2, vxc2, fxc2 = X.eval_eff(106, r, 1, "sigma")
    np.testing.assert_allclose(vxc2, xceff.transform(r, X.eval_raw(106, r, 1, 2), "sigma", 1), rtol=1e-15)
