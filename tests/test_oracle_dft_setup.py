"""Pin the DFT setup-pass oracle (fixed grid, RKS and UKS) with autograd derivatives of an
analytic model: the skeleton Hessian is d2 E_xc / dR dR at a fixed density matrix, drho is
d rho / dR, vmat_deriv1 is d V_xc / dR (V_xc = dE_xc / dD), and the vxc part of it is the
derivative with the potential frozen."""
import numpy as np
import pytest
import torch

from oracle import dft_setup as ods
from tests import dft_model as dm

DENS = ["rho", "sigma", "tau"]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def model():
    return dm.Model()


def test_component_table_matches_reference_constants():
    # hess_rks.rs:7-29
    assert [ods.comp(*c) for c in [(), (0,), (1,), (2,)]] == [0, 1, 2, 3]
    assert [ods.comp(0, 0), ods.comp(0, 1), ods.comp(0, 2), ods.comp(1, 1), ods.comp(1, 2), ods.comp(2, 2)] == \
        [4, 5, 6, 7, 8, 9]
    assert ods.comp(1, 0) == 5 and ods.comp(2, 1) == 8
    assert [ods.comp(0, 0, 0), ods.comp(0, 0, 1), ods.comp(0, 0, 2), ods.comp(0, 1, 1), ods.comp(0, 1, 2),
            ods.comp(0, 2, 2), ods.comp(1, 1, 1), ods.comp(1, 1, 2), ods.comp(1, 2, 2), ods.comp(2, 2, 2)] == \
        list(range(10, 20))


def test_model_ao_derivatives_against_autograd(model):
    """The closed-form AO derivatives of the model are the r-derivatives of its values."""
    ao = model.ao_np()
    g = torch.tensor(model.grid, requires_grad=True)
    saved = model.grid
    def val(gr):
        model.grid = gr
        return model.ao(model.R)[..., 0]
    try:
        for mu in range(model.nao):
            f = lambda gr: val(gr)[:, mu].sum()
            h1 = torch.autograd.functional.jacobian(f, g).numpy()
            np.testing.assert_allclose(ao[:, mu, 1:4], h1, rtol=1e-12, atol=1e-12)
            h2 = torch.autograd.functional.hessian(f, g)  # [ng, 3, ng, 3]
            d2 = np.einsum("gigj->gij", h2.numpy())
            for (i, j) in dm.PAIRS:
                np.testing.assert_allclose(ao[:, mu, ods.comp(i, j)], d2[:, i, j], rtol=1e-11, atol=1e-11)
    finally:
        model.grid = saved
    # third derivatives: finite differences of the closed-form second derivatives
    h = 1e-5
    for k in range(3):
        gp, gm = saved.copy(), saved.copy()
        gp[:, k] += h; gm[:, k] -= h
        model.grid = gp; ap = model.ao_np()
        model.grid = gm; am = model.ao_np()
        model.grid = saved
        for (i, j) in dm.PAIRS:
            fd = (ap[:, :, ods.comp(i, j)] - am[:, :, ods.comp(i, j)]) / (2 * h)
            np.testing.assert_allclose(ao[:, :, ods.comp(i, j, k)], fd, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("den", DENS)
def test_rho_and_drho(model, den):
    D = model.density_matrix(2, seed=1)
    ao = dm.ao_for(model, den)
    rho = ods.setup_rho(den, ao, D)
    ref = dm.rho_vars(model.ao(model.R), torch.as_tensor(D), den).numpy()
    np.testing.assert_allclose(rho, ref, rtol=1e-13, atol=1e-14)
    drho = ods.get_drho(den, ao, ods.ao_dm0(ao, D, den), model.aoslices)
    R = torch.tensor(model.R)
    jac = torch.autograd.functional.jacobian(lambda x: dm.rho_vars(model.ao(x), torch.as_tensor(D), den), R)
    np.testing.assert_allclose(drho, jac.numpy().transpose(0, 1, 3, 2), rtol=1e-11, atol=1e-12)


def sym(G):
    """The gradient with respect to a symmetric density matrix (the potential matrix)."""
    return 0.5 * (G + G.T)


def v_xc(model, R, D, den):
    Dt = torch.tensor(D, requires_grad=True)
    E = dm.energy(model, R, Dt, den)
    return sym(torch.autograd.grad(E, Dt, create_graph=True)[0])


@pytest.mark.parametrize("den", DENS)
def test_rks_setup_against_autograd(model, den):
    D = model.density_matrix(2, seed=2)
    ao = dm.ao_for(model, den)
    rho, vxc, fxc = dm.rks_inputs(model, D, den)
    r = ods.make_hessian_setup(den, ao, model.weights, D, model.aoslices, vxc, fxc)
    R = torch.tensor(model.R)
    H = torch.autograd.functional.hessian(lambda x: dm.energy(model, x, D, den), R).numpy()  # [A, t, B, s]
    assert rel(ods.skeleton_hess(r), H.transpose(1, 3, 0, 2)) < 1e-10
    # d V / dR at fixed D
    J = torch.autograd.functional.jacobian(lambda x: v_xc(model, x, D, den), R).numpy()  # [mu, nu, A, t]
    assert rel(r["vmat_deriv1"], J.transpose(0, 1, 3, 2)) < 1e-10
    # the vxc part: AO centres moved with the potential frozen at the reference density
    wv = torch.as_tensor(model.weights[:, None] * vxc)
    def v_frozen(x):
        Dt = torch.tensor(D, requires_grad=True)
        rr = dm.rho_vars(model.ao(x), Dt, den)
        return sym(torch.autograd.grad((wv * rr).sum(), Dt, create_graph=True)[0])
    Jf = torch.autograd.functional.jacobian(v_frozen, R).numpy()
    assert rel(ods.get_vmat_vxc(r["vmat_ip"], model.aoslices), Jf.transpose(0, 1, 3, 2)) < 1e-10


@pytest.mark.parametrize("den", DENS)
def test_rks_setup_batched_equals_unbatched(model, den):
    D = model.density_matrix(3, seed=3)
    ao = dm.ao_for(model, den)
    _, vxc, fxc = dm.rks_inputs(model, D, den)
    a = ods.make_hessian_setup(den, ao, model.weights, D, model.aoslices, vxc, fxc)
    b = ods.make_hessian_setup_batched(den, ao, model.weights, D, model.aoslices, vxc, fxc, nchunk=17)
    for k in a:
        np.testing.assert_allclose(b[k], a[k], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("den", DENS)
def test_uks_setup_against_autograd(model, den):
    Da, Db = model.density_matrix(2, seed=4, scale=1.0), model.density_matrix(1, seed=5, scale=1.0)
    ao = dm.ao_for(model, den)
    rho, vxc, fxc = dm.uks_inputs(model, Da, Db, den)
    np.testing.assert_allclose(ods.setup_rho_uks(den, ao, (Da, Db)), rho, rtol=1e-13, atol=1e-14)
    r = ods.make_hessian_setup_uks(den, ao, model.weights, (Da, Db), model.aoslices, vxc, fxc)
    R = torch.tensor(model.R)
    H = torch.autograd.functional.hessian(lambda x: dm.energy_uks(model, x, Da, Db, den), R).numpy()
    assert rel(ods.skeleton_hess_uks(r), H.transpose(1, 3, 0, 2)) < 1e-10
    for s, tag in enumerate("ab"):
        def v(x):
            D2 = [torch.tensor(Da), torch.tensor(Db)]
            D2[s] = D2[s].clone().requires_grad_(True)
            E = dm.energy_uks(model, x, D2[0], D2[1], den)
            return sym(torch.autograd.grad(E, D2[s], create_graph=True)[0])
        J = torch.autograd.functional.jacobian(v, R).numpy()
        assert rel(r["vmat_deriv1_" + tag], J.transpose(0, 1, 3, 2)) < 1e-10


@pytest.mark.parametrize("den", DENS)
def test_uks_closed_shell_equals_rks(model, den):
    """D_a = D_b = D/2 and eps_pol(ra, rb, ...) == eps_unpol at the spin-summed variables: the UKS
    skeleton Hessian equals the RKS one and each spin's vmat_deriv1 equals the RKS one."""
    D = model.density_matrix(2, seed=6)
    ao = dm.ao_for(model, den)
    _, vxc, fxc = dm.rks_inputs(model, D, den)
    r = ods.make_hessian_setup(den, ao, model.weights, D, model.aoslices, vxc, fxc)
    # the spin-polarised kernel of a functional that depends on the total density only
    nv = ods.NVAR[den]
    vp = np.stack([vxc, vxc], axis=2)
    fp = np.zeros((len(model.weights), nv, 2, nv, 2))
    for s1 in range(2):
        for s2 in range(2):
            fp[:, :, s1, :, s2] = fxc
    u = ods.make_hessian_setup_uks(den, ao, model.weights, (D / 2, D / 2), model.aoslices, vp, fp)
    assert rel(ods.skeleton_hess_uks(u), ods.skeleton_hess(r)) < 1e-12
    for tag in "ab":
        assert rel(u["vmat_deriv1_" + tag], r["vmat_deriv1"]) < 1e-12
