"""Pin the numint and xceff oracles independently of their own formulas (CPU)."""
import numpy as np
import pytest
import torch

from oracle import numint as oni
from oracle import xceff as oxc


def grid_case(ng=17, nao=6, nocc=2, nset=3, seed=0, den="tau"):
    rng = np.random.default_rng(seed)
    ao = rng.standard_normal((ng, nao, 4))
    w = rng.uniform(0.1, 1.0, ng)
    bra = rng.standard_normal((nao, nocc))
    kets = rng.standard_normal((nao, nocc, nset))
    nv = oni.NVAR[den]
    f = rng.standard_normal((ng, nv, nv))
    f = f + f.transpose(0, 2, 1)
    return ao, w, bra, kets, f


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
def test_rho1_gemm_form_equals_pointwise_loops(den):
    ao, w, bra, kets, f = grid_case(den=den)
    np.testing.assert_allclose(oni.rho1_one_bra_mult_ket(ao, bra, kets, den), oni.brute_rho1(ao, bra, kets, den),
                               rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
def test_bra_transformed_potential_equals_full_potential_times_bra(den):
    ao, w, bra, kets, f = grid_case(den=den, seed=1)
    rho1 = oni.rho1_one_bra_mult_ket(ao, bra, kets, den)
    V = oni.brute_fxc_full(den, f, rho1, ao, w)
    out = oni.fxc_pot_bra_trans(den, f, rho1, ao, w, bra, nchunk=5)
    np.testing.assert_allclose(out, np.einsum("uva,vi->uia", V, bra), rtol=1e-12, atol=1e-12)


def energy_density(rho, sigma, tau=None):
    """A smooth analytic meta-GGA-like energy density (not a physical functional)."""
    e = -0.7 * rho ** (4.0 / 3.0) + 0.05 * sigma / (1.0 + rho) + 0.01 * torch.log1p(sigma) * rho
    if tau is not None:
        e = e + 0.03 * tau * rho / (1.0 + tau) + 0.002 * tau * sigma
    return e


@pytest.mark.parametrize("den", ["sigma", "tau"])
def test_xceff_chain_rule_against_autograd(den):
    """fxc_eff from the LibXC-variable derivatives == Hessian of e(rho, |grad rho|^2, tau) in rho_t."""
    rng = np.random.default_rng(2)
    ng = 7
    nv = oxc.NVAR[den]
    xlen = oxc.XLEN[den]
    rho = np.empty((ng, nv))
    rho[:, 0] = rng.uniform(0.2, 1.5, ng)
    rho[:, 1:4] = rng.standard_normal((ng, 3)) * 0.4
    if den == "tau":
        rho[:, 4] = rng.uniform(0.1, 1.0, ng)
    xc = np.zeros((ng, 1 + xlen + xlen * (xlen + 1) // 2))
    vxc_ref = np.zeros((ng, nv))
    fxc_ref = np.zeros((ng, nv, nv))
    for g in range(ng):
        # derivatives in LibXC variables (rho, sigma, tau)
        v = torch.tensor([rho[g, 0], float(rho[g, 1:4] @ rho[g, 1:4])] + ([rho[g, 4]] if den == "tau" else []),
                         dtype=torch.float64, requires_grad=True)
        f = lambda q: energy_density(q[0], q[1], q[2] if den == "tau" else None)
        d1 = torch.autograd.functional.jacobian(f, v).numpy()
        d2 = torch.autograd.functional.hessian(f, v).numpy()
        xc[g, 0] = f(v).item()
        xc[g, 1:1 + xlen] = d1
        for k, (p, q) in enumerate(oxc.uniq_pairs(xlen)):
            xc[g, 1 + xlen + k] = d2[p, q]
        # derivatives in rho_t variables (rho, grad_x, grad_y, grad_z, tau)
        u = torch.tensor(rho[g, :nv], dtype=torch.float64, requires_grad=True)
        h = lambda r: energy_density(r[0], r[1:4] @ r[1:4], r[4] if den == "tau" else None)
        vxc_ref[g] = torch.autograd.functional.jacobian(h, u).numpy()
        fxc_ref[g] = torch.autograd.functional.hessian(h, u).numpy()
    np.testing.assert_allclose(oxc.transform(rho, xc, den, 1), vxc_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oxc.transform(rho, xc, den, 2), fxc_ref, rtol=1e-11, atol=1e-12)
