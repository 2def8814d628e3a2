"""Pin the RI-JK skeleton oracle (the reference's naive term list) with autograd derivatives of
an analytic integral model: the J / K skeleton Hessians are the second derivatives of
E_J = 1/2 (D|P) V^-1 (P|D) and E_K = 1/2 sum_ij (ij|P) V^-1 (P|ij) at fixed orbitals, and the
first-derivative Fock terms are dJ/dR and dK/dR."""
import numpy as np
import pytest
import torch

from oracle import rijk_skeleton as osk
from tests.rijk_model import Model


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def model():
    return Model()


@pytest.fixture(scope="module")
def ints(model):
    return model.integrals()


def test_model_derivative_integrals_against_autograd(model, ints):
    from torch.func import jacfwd, jacrev
    R = model.R
    cu = torch.tensor(R[model.ao_atom]); cP = torch.tensor(R[model.aux_atom])
    f3, f2 = model.j3c_centres, model.j2c_centres
    npy = lambda t: t.detach().numpy()
    # first derivatives: minus the derivative with respect to the centre of the function
    ip1 = -np.einsum("uvPuc->uvPc", npy(jacrev(lambda x: f3(x, cu, cP))(cu)))
    ip2 = -np.einsum("uvPPc->uvPc", npy(jacrev(lambda x: f3(cu, cu, x))(cP)))
    np.testing.assert_allclose(ints["ip1"], ip1, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(ints["ip2"], ip2, rtol=1e-12, atol=1e-13)
    # second derivatives [.., a, b]: a on the later function, b on the earlier one
    Huu = npy(jacfwd(jacrev(lambda x: f3(x, cu, cP)))(cu))                         # [u,v,P, u',a, u'',b]
    np.testing.assert_allclose(ints["ipip1"], np.einsum("uvPuaub->uvPab", Huu), rtol=1e-11, atol=1e-12)
    Huv = npy(jacfwd(lambda y: jacrev(lambda x: f3(y, x, cP))(cu))(cu))            # [u,v,P, v',a, u',b]
    np.testing.assert_allclose(ints["ipvip1"], np.einsum("uvPvaub->uvPab", Huv), rtol=1e-11, atol=1e-12)
    HuP = npy(jacfwd(lambda y: jacrev(lambda x: f3(y, cu, x))(cP))(cu))            # [u,v,P, P',a, u',b]
    np.testing.assert_allclose(ints["ip1ip2"], np.einsum("uvPPaub->uvPab", HuP), rtol=1e-11, atol=1e-12)
    HPP = npy(jacfwd(jacrev(lambda x: f3(cu, cu, x)))(cP))
    np.testing.assert_allclose(ints["ipip2"], np.einsum("uvPPaPb->uvPab", HPP), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(ints["j2c_ip1"], -np.einsum("PQPc->PQc", npy(jacrev(lambda x: f2(x, cP))(cP))),
                               rtol=1e-12, atol=1e-13)
    H2 = npy(jacfwd(jacrev(lambda x: f2(x, cP)))(cP))
    np.testing.assert_allclose(ints["j2c_ipip1"], np.einsum("PQPaPb->PQab", H2), rtol=1e-11, atol=1e-12)
    H2b = npy(jacfwd(lambda y: jacrev(lambda x: f2(y, x))(cP))(cP))                # [P,Q, Q',a, P',b]
    np.testing.assert_allclose(ints["j2c_ip1ip2"], np.einsum("PQQaPb->PQab", H2b), rtol=1e-11, atol=1e-12)


def hess(f, R):
    H = torch.autograd.functional.hessian(f, torch.tensor(R)).numpy()   # [A, t, B, s]
    return H.transpose(1, 3, 0, 2)


def test_j_skeleton_is_the_hessian_of_the_coulomb_energy(model, ints):
    C, occ = model.orbitals(3, seed=1)
    mocc2 = C[:, :3] * np.sqrt(2.0)
    D = mocc2 @ mocc2.T
    de = sum(osk.rij_skeleton_naive(ints, D, model.aoslices, model.auxslices).values())
    assert rel(de, hess(lambda x: model.e_j(x, D), model.R)) < 1e-10


def test_k_skeleton_is_the_hessian_of_the_exchange_energy(model, ints):
    C, occ = model.orbitals(3, seed=2)
    mocc2 = C[:, :3] * np.sqrt(2.0)
    de = sum(osk.rik_skeleton_naive(ints, mocc2, model.aoslices, model.auxslices).values())
    assert rel(de, hess(lambda x: model.e_k(x, mocc2), model.R)) < 1e-10


def test_deriv1_fock_terms_against_autograd(model, ints):
    C, occ = model.orbitals(3, seed=3)
    mocc2 = C[:, :3] * np.sqrt(2.0)
    D = mocc2 @ mocc2.T
    R = torch.tensor(model.R)
    jj = torch.autograd.functional.jacobian(lambda x: model.fock_j(x, D), R).numpy()   # [u, v, A, t]
    j1 = osk.rij_deriv1_ao_naive(ints, D, model.aoslices, model.auxslices)
    assert rel(j1, jj.transpose(0, 1, 3, 2)) < 1e-10
    kk = torch.autograd.functional.jacobian(lambda x: model.fock_k(x, mocc2), R).numpy()
    k1 = osk.rik_deriv1_ao_naive(ints, mocc2, model.aoslices, model.auxslices)
    assert rel(k1, kk.transpose(0, 1, 3, 2)) < 1e-10


def test_translational_invariance_and_closed_shell_uhf(model, ints):
    C, occ = model.orbitals(3, seed=4)
    de, d1 = osk.skeleton_rhf(ints, C, occ, 1.0, 0.6, model.aoslices, model.auxslices)
    # the model integrals are translation invariant: rows of the skeleton Hessian sum to zero
    assert np.abs(de.sum(axis=3)).max() < 1e-10 * np.abs(de).max()
    occ1 = np.where(occ > 0, 1.0, 0.0)
    deu, d1u = osk.skeleton_uhf(ints, [C, C], [occ1, occ1], 1.0, 0.6, model.aoslices, model.auxslices)
    assert rel(deu, de) < 1e-12
    for s in range(2):
        assert rel(d1u[s], d1) < 1e-12
