"""GPU parity: DFT skeleton-Hessian setup passes (RKS and UKS, fixed grid) against the oracle,
on the analytic model (whose oracle results equal autograd derivatives) and on larger random
AO tensors; grid batching and atom lists.  FP64: 1e-11 relative to max|ref|."""
import numpy as np
import pytest

from oracle import dft_setup as ods
from tests import dft_model as dm
from tests.test_gpu_cphf import T, rel

pytestmark = pytest.mark.gpu

DENS = ["rho", "sigma", "tau"]


def check(got, ref, tol=1e-11):
    for k, v in ref.items():
        e = rel(got[k].cpu().numpy(), v)
        assert e < tol, (k, e)


@pytest.mark.parametrize("den", DENS)
@pytest.mark.parametrize("nbatch", [0, 64])
def test_rks_setup_model(dev, den, nbatch):
    import hesship
    model = dm.Model(ng=300, seed=1)
    ao = dm.ao_for(model, den)
    D = model.density_matrix(2, seed=2)
    rho, vxc, fxc = dm.rks_inputs(model, D, den)
    ni = hesship.NIMatmul(T(ao, dev), T(model.weights, dev), nbatch=nbatch)
    assert rel(ni.setup_rho(den, T(D, dev)).cpu().numpy(), ods.setup_rho(den, ao, D)) < 1e-12
    got = ni.make_hessian_setup(den, T(D, dev), model.aoslices, T(vxc, dev), T(fxc, dev))
    ref = ods.make_hessian_setup(den, ao, model.weights, D, model.aoslices, vxc, fxc)
    check(got, ref)
    assert rel(ni.skeleton_hess(got).cpu().numpy(), ods.skeleton_hess(ref)) < 1e-11


@pytest.mark.parametrize("den", DENS)
def test_uks_setup_model(dev, den):
    import hesship
    model = dm.Model(ng=250, seed=3)
    ao = dm.ao_for(model, den)
    Da, Db = model.density_matrix(2, seed=4, scale=1.0), model.density_matrix(1, seed=5, scale=1.0)
    rho, vxc, fxc = dm.uks_inputs(model, Da, Db, den)
    ni = hesship.NIMatmul(T(ao, dev), T(model.weights, dev), nbatch=128)
    assert rel(ni.setup_rho(den, [T(Da, dev), T(Db, dev)]).cpu().numpy(), rho) < 1e-12
    got = ni.make_hessian_setup(den, [T(Da, dev), T(Db, dev)], model.aoslices, T(vxc, dev), T(fxc, dev))
    ref = ods.make_hessian_setup_uks(den, ao, model.weights, (Da, Db), model.aoslices, vxc, fxc)
    check(got, ref)
    assert rel(ni.skeleton_hess(got).cpu().numpy(), ods.skeleton_hess_uks(ref)) < 1e-11


def random_case(ng, nao, nbas, den, seed, nspin=1):
    rng = np.random.default_rng(seed)
    ao = rng.standard_normal((ng, nao, ods.NCOMP_AO[ods.AO_DERIV[den]])) * 0.5
    w = rng.uniform(0.1, 1.0, ng)
    nv = ods.NVAR[den]
    C = [rng.standard_normal((nao, 5)) * 0.3 for _ in range(nspin)]
    dm0 = [c @ c.T for c in C]
    vxc = rng.standard_normal((ng, nv, nspin))
    f = rng.standard_normal((ng, nv, nspin, nv, nspin))
    f = f + f.transpose(0, 3, 4, 1, 2)
    off = np.concatenate([[0], np.cumsum(nbas)])
    slices = [(int(off[A]), int(off[A + 1])) for A in range(len(nbas))]
    return ao, w, dm0, vxc, f, slices


@pytest.mark.parametrize("den", DENS)
def test_rks_setup_random_atom_list(dev, den):
    """Larger random tensors; the atom list is a reordered subset (aoslice_by_atom()[atm_list])."""
    import hesship
    ao, w, dm0, vxc, f, slices = random_case(2000, 46, (7, 9, 11, 6, 13), den, seed=6)
    sub = [slices[3], slices[0], slices[4]]
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=704)
    got = ni.make_hessian_setup(den, T(dm0[0], dev), sub, T(vxc[:, :, 0], dev), T(f[:, :, 0, :, 0], dev))
    ref = ods.make_hessian_setup(den, ao, w, dm0[0], sub, vxc[:, :, 0], f[:, :, 0, :, 0])
    check(got, ref)


@pytest.mark.parametrize("den", ["sigma", "tau"])
def test_uks_setup_random(dev, den):
    import hesship
    ao, w, dm0, vxc, f, slices = random_case(1500, 38, (10, 12, 16), den, seed=7, nspin=2)
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev))
    got = ni.make_hessian_setup(den, [T(d, dev) for d in dm0], slices, T(vxc, dev), T(f, dev))
    ref = ods.make_hessian_setup_uks(den, ao, w, tuple(dm0), slices, vxc, f)
    check(got, ref)


def test_setup_needs_ao_derivatives(dev):
    import hesship
    ao, w, dm0, vxc, f, slices = random_case(100, 10, (4, 6), "sigma", seed=8)
    ni = hesship.NIMatmul(T(ao[:, :, :10], dev), T(w, dev))
    with pytest.raises(hesship.HesshipError, match="derivatives"):
        ni.make_hessian_setup("sigma", T(dm0[0], dev), slices, T(vxc[:, :, 0], dev), T(f[:, :, 0, :, 0], dev))
    with pytest.raises(hesship.HesshipError, match="overlap"):
        ni.make_hessian_setup("rho", T(dm0[0], dev), [(0, 5), (4, 10)], T(vxc[:, :1, 0], dev),
                              T(f[:, :1, 0, :1, 0], dev))
