"""Pin the Becke-partition oracle and the Becke-augmented DFT setup oracle.

Partition (becke_partition.rs): the switch derivatives against autograd for several hardness
values; w against an explicit per-pair loop and the partition of unity; dw / ddw against
autograd of the directly written weights with the grid carried by its atom; translation
invariance and the (A, t) <-> (B, s) symmetry; ByAtom == ByGrid; batch-size independence; the
c / dc / ddc contractions; padding grids.

Setup (hess_rks_becke.rs / hess_uks_becke.rs): de_xc_skeleton is d2 E_xc / dR dR of the XC
energy on the glued, Becke-weighted grid at a fixed density matrix; vmat_deriv1_grid is d V_xc
/ dR of the same energy; both are translation invariant; the grid-fixed keys equal the plain
setup's; chunking does not change the result; UKS with equal spins equals RKS.
"""
import numpy as np
import pytest
import torch

from oracle import becke as obk
from oracle import dft_becke as odb
from oracle import dft_setup as ods
from tests import becke_model as bm
from tests import dft_model as dm

DENS = ["rho", "sigma", "tau"]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def model():
    return bm.BeckeModel()


def run(model, deriv=2, hardness=None, **kw):
    return obk.partition(model.grid, model.R, model.attr, model.wquad, model.adjust,
                         hardness or model.hardness, deriv, **kw)


@pytest.mark.parametrize("h", [1, 2, 3, 4, 5])
def test_switch_derivatives_against_autograd(h):
    mu = torch.linspace(-0.97, 0.97, 41, dtype=torch.float64)
    a = 0.27
    def f(m):
        x = m + a * (1 - m * m)
        for _ in range(h):
            x = 1.5 * x - 0.5 * x ** 3
        return x
    f0, f1, f2 = obk.switch(mu.numpy(), a, h)
    # derivatives with respect to nu: divide the mu-derivatives by dnu/dmu (and its derivative)
    m = mu.clone().requires_grad_(True)
    y = f(m)
    g1 = torch.autograd.grad(y.sum(), m, create_graph=True)[0]
    g2 = torch.autograd.grad(g1.sum(), m)[0]
    dnu = (1 - 2 * a * mu).numpy()
    np.testing.assert_allclose(f0, y.detach().numpy(), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(f1 * dnu, g1.detach().numpy(), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(f2 * dnu ** 2 + f1 * (-2 * a), g2.numpy(), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("h", [1, 2, 3, 4])
def test_weights_against_pair_loop(model, h):
    w = run(model, 0, hardness=h)["w"]
    ref = obk.partition_weights_loop(model.grid, model.R, model.attr, model.wquad, model.adjust, h)
    np.testing.assert_allclose(w, ref, rtol=1e-13, atol=1e-15)


def test_partition_of_unity(model):
    """Every grid point attributed to every atom in turn: the cell weights sum to wquad."""
    natm, ng = model.natm, len(model.attr)
    coords = np.tile(model.grid, (natm, 1))
    attr = np.repeat(np.arange(natm), ng)
    wq = np.tile(model.wquad, natm)
    w = obk.partition(coords, model.R, attr, wq, model.adjust, 3, 0)["w"]
    np.testing.assert_allclose(w.reshape(natm, ng).sum(axis=0), model.wquad, rtol=1e-13)


@pytest.mark.parametrize("h", [3, 2])
def test_dw_ddw_against_autograd(h):
    m = bm.BeckeModel(hardness=h, seed=3)
    r = run(m, 2)
    R = torch.tensor(m.R)
    J = torch.autograd.functional.jacobian(m.becke_w, R).numpy()          # [g, A, t]
    np.testing.assert_allclose(r["w"], m.weights, rtol=1e-12, atol=1e-15)
    assert rel(r["dw"], J.transpose(1, 2, 0)) < 1e-11
    Hs = [torch.autograd.functional.hessian(lambda x: m.becke_w(x)[g], R).numpy() for g in range(len(m.attr))]
    H = np.stack(Hs, axis=-1)                                              # [A, t, B, s, g]
    assert rel(r["ddw"], H) < 1e-9


def test_invariances(model):
    r = run(model, 2)
    dw, ddw = r["dw"], r["ddw"]
    s = np.abs(dw).max()
    assert np.abs(dw.sum(axis=0)).max() < 1e-13 * s
    s2 = np.abs(ddw).max()
    assert np.abs(ddw.sum(axis=0)).max() < 1e-12 * s2
    assert np.abs(ddw.sum(axis=2)).max() < 1e-12 * s2
    np.testing.assert_allclose(ddw, ddw.transpose(2, 3, 0, 1, 4), rtol=1e-12, atol=1e-12 * s2)


def test_batching_contractions_and_padding(model):
    a = run(model, 2, nbatch=7)
    b = run(model, 2, nbatch=1000)
    for k in ("w", "dw", "ddw"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-14, atol=1e-16)
    rng = np.random.default_rng(9)
    ng = len(model.attr)
    cw = rng.standard_normal((2, ng))
    c = run(model, 2, contract_w=cw, contract_dw=cw, contract_ddw=cw, output=(False, False, False))
    assert set(c) == {"c", "dc", "ddc"}
    np.testing.assert_allclose(c["c"], cw @ a["w"], rtol=1e-12)
    np.testing.assert_allclose(c["dc"], np.einsum("atg,kg->atk", a["dw"], cw), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(c["ddc"], np.einsum("atbsg,kg->atbsk", a["ddw"], cw), rtol=1e-12, atol=1e-13)
    # padding attribution carries nothing; the other points are unchanged
    attr = model.attr.copy()
    attr[::5] = model.natm + 2
    p = obk.partition(model.grid, model.R, attr, model.wquad, model.adjust, 3, 2)
    pad = attr >= model.natm
    assert not p["w"][pad].any() and not p["dw"][..., pad].any() and not p["ddw"][..., pad].any()
    np.testing.assert_array_equal(p["w"][~pad], a["w"][~pad])
    np.testing.assert_array_equal(p["dw"][..., ~pad], a["dw"][..., ~pad])
    # ByAtom boundaries reproduce the per-grid attribution
    np.testing.assert_array_equal(obk.by_atom_attribution(model.split), model.attr)


# ---------------------------------------------------------------------------------------------
# setup with the grid shift

def sym(G):
    return 0.5 * (G + G.T)


def becke_inputs(model, D, den):
    ao = dm.ao_for(model, den)
    _, vxc, fxc = dm.rks_inputs(model, D, den)
    return ao, vxc, fxc, bm.exc_rks(model, D, den)


def setup(model, den, D, nchunk=None):
    ao, vxc, fxc, exc = becke_inputs(model, D, den)
    return odb.make_hessian_setup_becke(den, ao, model.weights, model.wquad, model.grid, model.R, model.adjust,
                                        model.hardness, D, model.aoslices, model.split, exc, vxc, fxc, nchunk)


@pytest.mark.parametrize("den", DENS)
def test_rks_becke_setup_against_autograd(model, den):
    D = model.density_matrix(2, seed=2)
    r = setup(model, den, D)
    R = torch.tensor(model.R)
    H = torch.autograd.functional.hessian(lambda x: bm.energy(model, x, D, den), R).numpy()
    assert rel(r["de_xc_skeleton"], H.transpose(1, 3, 0, 2)) < 1e-10
    def v(x):
        Dt = torch.tensor(D, requires_grad=True)
        return sym(torch.autograd.grad(bm.energy(model, x, Dt, den), Dt, create_graph=True)[0])
    J = torch.autograd.functional.jacobian(v, R).numpy()
    assert rel(r["vmat_deriv1_grid"], J.transpose(0, 1, 3, 2)) < 1e-10
    # translation invariance of both, and the grid-fixed keys of the plain setup
    s = np.abs(r["de_xc_skeleton"]).max()
    assert np.abs(r["de_xc_skeleton"].sum(axis=3)).max() < 1e-11 * s
    assert np.abs(r["vmat_deriv1_grid"].sum(axis=3)).max() < 1e-11 * np.abs(r["vmat_deriv1_grid"]).max()
    ao, vxc, fxc, _ = becke_inputs(model, D, den)
    plain = ods.make_hessian_setup(den, ao, model.weights, D, model.aoslices, vxc, fxc)
    for k in plain:
        np.testing.assert_allclose(r[k], plain[k], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("den", DENS)
def test_rks_becke_chunked_equals_unchunked(model, den):
    D = model.density_matrix(3, seed=3)
    a = setup(model, den, D)
    b = setup(model, den, D, nchunk=7)
    for k in a:
        np.testing.assert_allclose(b[k], a[k], rtol=1e-11, atol=1e-13 * max(1.0, np.abs(a[k]).max()))


@pytest.mark.parametrize("den", DENS)
def test_uks_becke_setup_against_autograd(model, den):
    Da, Db = model.density_matrix(2, seed=4, scale=1.0), model.density_matrix(1, seed=5, scale=1.0)
    ao = dm.ao_for(model, den)
    _, vxc, fxc = dm.uks_inputs(model, Da, Db, den)
    exc = bm.exc_uks(model, Da, Db, den)
    r = odb.make_hessian_setup_becke_uks(den, ao, model.weights, model.wquad, model.grid, model.R, model.adjust,
                                         model.hardness, (Da, Db), model.aoslices, model.split, exc, vxc, fxc)
    R = torch.tensor(model.R)
    H = torch.autograd.functional.hessian(lambda x: bm.energy_uks(model, x, Da, Db, den), R).numpy()
    assert rel(r["de_xc_skeleton"], H.transpose(1, 3, 0, 2)) < 1e-10
    for s, tag in enumerate("ab"):
        def v(x):
            D2 = [torch.tensor(Da), torch.tensor(Db)]
            D2[s] = D2[s].clone().requires_grad_(True)
            E = bm.energy_uks(model, x, D2[0], D2[1], den)
            return sym(torch.autograd.grad(E, D2[s], create_graph=True)[0])
        J = torch.autograd.functional.jacobian(v, R).numpy()
        assert rel(r["vmat_deriv1_grid_" + tag], J.transpose(0, 1, 3, 2)) < 1e-10


@pytest.mark.parametrize("den", DENS)
def test_uks_becke_closed_shell_equals_rks(model, den):
    D = model.density_matrix(2, seed=6)
    ao, vxc, fxc, exc = becke_inputs(model, D, den)
    r = odb.make_hessian_setup_becke(den, ao, model.weights, model.wquad, model.grid, model.R, model.adjust,
                                     model.hardness, D, model.aoslices, model.split, exc, vxc, fxc)
    nv = ods.NVAR[den]
    vp = np.stack([vxc, vxc], axis=2)
    fp = np.zeros((len(model.weights), nv, 2, nv, 2))
    for s1 in range(2):
        for s2 in range(2):
            fp[:, :, s1, :, s2] = fxc
    u = odb.make_hessian_setup_becke_uks(den, ao, model.weights, model.wquad, model.grid, model.R, model.adjust,
                                         model.hardness, (D / 2, D / 2), model.aoslices, model.split, exc, vp, fp)
    assert rel(u["de_xc_skeleton"], r["de_xc_skeleton"]) < 1e-12
    for tag in "ab":
        assert rel(u["vmat_deriv1_grid_" + tag], r["vmat_deriv1_grid"]) < 1e-12
