"""GPU parity: Becke partition (w, dw, ddw and the c / dc / ddc contractions; ByAtom and ByGrid
with padding; hardness 2 and 3) and the DFT setup with the Becke grid shift (RKS and UKS; the
analytic model whose oracle results equal autograd derivatives, several batch sizes) against the
oracle.  FP64: 1e-11 relative to max|ref| (1e-10 for the second derivatives)."""
import numpy as np
import pytest

from oracle import becke as obk
from oracle import dft_becke as odb
from tests import becke_model as bm
from tests import dft_model as dm
from tests.test_gpu_cphf import T, rel

pytestmark = pytest.mark.gpu

DENS = ["rho", "sigma", "tau"]


def geometry(natm, npa, seed):
    rng = np.random.default_rng(seed)
    R = rng.uniform(-2.0, 2.0, (natm, 3))
    attr = np.repeat(np.arange(natm), npa)
    u = rng.standard_normal((len(attr), 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    coords = R[attr] + u * rng.uniform(0.1, 3.0, (len(attr), 1))
    wq = rng.uniform(0.01, 0.3, len(attr))
    a = rng.uniform(-0.4, 0.4, (natm, natm))
    adj = np.triu(a, 1) - np.triu(a, 1).T
    split = np.concatenate([[0], np.cumsum(npa)]).astype(np.int64)
    return R, attr, coords, wq, adj, split


def close(got, ref, tol):
    e = rel(got.cpu().numpy(), ref)
    assert e < tol, e


@pytest.mark.parametrize("h", [3, 2])
def test_partition_by_atom(dev, h):
    import hesship
    R, attr, coords, wq, adj, split = geometry(5, (37, 0, 52, 23, 41), seed=h)
    tab = hesship.BeckeMolTables(R, adj, dev)
    rng = np.random.default_rng(7)
    cw = rng.standard_normal((2, len(attr)))
    got = tab.partition(T(coords, dev), split, T(wq, dev), hardness=h, deriv=2, by_atom=True,
                        contract_w=T(cw, dev), contract_dw=T(cw, dev), contract_ddw=T(cw, dev))
    ref = obk.partition(coords, R, attr, wq, adj, h, 2, contract_w=cw, contract_dw=cw, contract_ddw=cw)
    close(got["w"], ref["w"], 1e-12)
    close(got["dw"], ref["dw"], 1e-11)
    close(got["ddw"], ref["ddw"], 1e-10)
    close(got["c"], ref["c"], 1e-12)
    close(got["dc"], ref["dc"], 1e-11)
    close(got["ddc"], ref["ddc"], 1e-10)


def test_partition_by_grid_padding_and_orders(dev):
    """Interleaved attribution with padding points; deriv 0 / 1 outputs; outputs not requested
    are absent; dw alone equals dw of the deriv-2 call."""
    import hesship
    R, attr, coords, wq, adj, split = geometry(4, (30, 25, 40, 35), seed=11)
    rng = np.random.default_rng(3)
    order = rng.permutation(len(attr))
    attr, coords, wq = attr[order].copy(), coords[order], wq[order]
    attr[::7] = 4 + (np.arange(len(attr[::7])) % 3)
    tab = hesship.BeckeMolTables(R, adj, dev)
    ref = obk.partition(coords, R, attr, wq, adj, 3, 2)
    got = tab.partition(T(coords, dev), attr, T(wq, dev), deriv=2)
    for k, tol in (("w", 1e-12), ("dw", 1e-11), ("ddw", 1e-10)):
        close(got[k], ref[k], tol)
    g1 = tab.partition(T(coords, dev), attr, T(wq, dev), deriv=1, output=(False, True, True))
    assert set(g1) == {"dw"}
    close(g1["dw"], ref["dw"], 1e-11)
    g0 = tab.partition(T(coords, dev), attr, T(wq, dev), deriv=0)
    assert set(g0) == {"w"}
    close(g0["w"], ref["w"], 1e-12)


def test_partition_checks(dev):
    import hesship
    R, attr, coords, wq, adj, split = geometry(3, (10, 12, 9), seed=5)
    tab = hesship.BeckeMolTables(R, adj, dev)
    bad = split.copy()
    bad[1] = 40
    with pytest.raises(hesship.HesshipError, match="must not exceed ngrids|non-decreasing"):
        tab.partition(T(coords, dev), bad, T(wq, dev), by_atom=True)
    bad = split.copy()
    bad[-1] -= 1
    with pytest.raises(hesship.HesshipError, match="must end with ngrids"):
        tab.partition(T(coords, dev), bad, T(wq, dev), by_atom=True)
    with pytest.raises(hesship.HesshipError, match="deriv"):
        tab.partition(T(coords, dev), split, T(wq, dev), deriv=3, by_atom=True)
    R50, attr50, coords50, wq50, adj50, split50 = geometry(50, (3,) * 50, seed=9)
    R50 *= 3.0
    coords50 = R50[attr50] + (coords50 - geometry(50, (3,) * 50, seed=9)[0][attr50])
    big = hesship.BeckeMolTables(R50, adj50, dev)
    with pytest.raises(hesship.HesshipError, match="48 atoms"):
        big.partition(T(coords50, dev), split50, T(wq50, dev), deriv=2, by_atom=True)
    # deriv 1 has no such limit
    g = big.partition(T(coords50, dev), split50, T(wq50, dev), deriv=1, by_atom=True)
    ref = obk.partition(coords50, R50, attr50, wq50, adj50, 3, 1)
    close(g["w"], ref["w"], 1e-12)
    close(g["dw"], ref["dw"], 1e-10)


# ---------------------------------------------------------------------------------------------
# setup with the grid shift

def model_case(den, seed, npa=(41, 29, 53)):
    model = bm.BeckeModel(npa=npa, seed=seed)
    return model, dm.ao_for(model, den)


def run_rks(dev, model, ao, den, D, nbatch):
    import hesship
    _, vxc, fxc = dm.rks_inputs(model, D, den)
    exc = bm.exc_rks(model, D, den)
    ni = hesship.NIMatmul(T(ao, dev), T(model.weights, dev), nbatch=nbatch)
    tab = hesship.BeckeMolTables(model.R, model.adjust, dev)
    got = ni.make_hessian_setup_becke(den, T(D, dev), model.aoslices, T(exc, dev), T(vxc, dev), T(fxc, dev), tab,
                                      T(model.grid, dev), T(model.wquad, dev), model.split, model.hardness)
    ref = odb.make_hessian_setup_becke(den, ao, model.weights, model.wquad, model.grid, model.R, model.adjust,
                                       model.hardness, D, model.aoslices, model.split, exc, vxc, fxc)
    return got, ref


@pytest.mark.parametrize("den", DENS)
@pytest.mark.parametrize("nbatch", [0, 64])
def test_rks_becke_setup_model(dev, den, nbatch):
    model, ao = model_case(den, seed=1)
    D = model.density_matrix(2, seed=2)
    got, ref = run_rks(dev, model, ao, den, D, nbatch)
    for k in ("de_xc_skeleton", "vmat_deriv1_grid", "de_fxc", "de_vxc_diag", "de_vxc_off", "vmat_ip", "vmat_deriv1",
              "de_xc_skeleton_no_becke"):
        e = rel(got[k].cpu().numpy(), ref[k])
        assert e < 1e-10, (k, e)
    # translation invariance of the device results
    de = got["de_xc_skeleton"].cpu().numpy()
    assert np.abs(de.sum(axis=3)).max() < 1e-10 * np.abs(de).max()


@pytest.mark.parametrize("den", DENS)
def test_uks_becke_setup_model(dev, den):
    import hesship
    model, ao = model_case(den, seed=4, npa=(35, 44, 30))
    Da, Db = model.density_matrix(2, seed=4, scale=1.0), model.density_matrix(1, seed=5, scale=1.0)
    _, vxc, fxc = dm.uks_inputs(model, Da, Db, den)
    exc = bm.exc_uks(model, Da, Db, den)
    ni = hesship.NIMatmul(T(ao, dev), T(model.weights, dev), nbatch=64)
    tab = hesship.BeckeMolTables(model.R, model.adjust, dev)
    got = ni.make_hessian_setup_becke(den, [T(Da, dev), T(Db, dev)], model.aoslices, T(exc, dev), T(vxc, dev),
                                      T(fxc, dev), tab, T(model.grid, dev), T(model.wquad, dev), model.split)
    ref = odb.make_hessian_setup_becke_uks(den, ao, model.weights, model.wquad, model.grid, model.R, model.adjust,
                                           model.hardness, (Da, Db), model.aoslices, model.split, exc, vxc, fxc)
    for k in ("de_xc_skeleton", "vmat_deriv1_grid_a", "vmat_deriv1_grid_b", "de_fxc", "de_vxc_diag_a",
              "de_vxc_off_b", "vmat_deriv1_a", "vmat_ip_b", "de_xc_skeleton_no_becke"):
        e = rel(got[k].cpu().numpy(), ref[k])
        assert e < 1e-10, (k, e)


def test_becke_setup_grid_keys_only_and_checks(dev):
    """Without the grid-fixed keys (the consumer's call) the two main results are unchanged;
    a partial atom list or a foreign atom count is rejected."""
    import hesship
    den = "sigma"
    model, ao = model_case(den, seed=6)
    D = model.density_matrix(2, seed=3)
    _, vxc, fxc = dm.rks_inputs(model, D, den)
    exc = bm.exc_rks(model, D, den)
    ni = hesship.NIMatmul(T(ao, dev), T(model.weights, dev), nbatch=128)
    tab = hesship.BeckeMolTables(model.R, model.adjust, dev)
    args = (T(exc, dev), T(vxc, dev), T(fxc, dev), tab, T(model.grid, dev), T(model.wquad, dev))
    full = ni.make_hessian_setup_becke(den, T(D, dev), model.aoslices, *args, model.split)
    lean = ni.make_hessian_setup_becke(den, T(D, dev), model.aoslices, *args, model.split, grid_fixed_keys=False)
    assert set(lean) == {"de_xc_skeleton", "vmat_deriv1_grid"}
    for k in lean:
        assert rel(lean[k].cpu().numpy(), full[k].cpu().numpy()) < 1e-13
    with pytest.raises(hesship.HesshipError, match="full atom list"):
        ni.make_hessian_setup_becke(den, T(D, dev), model.aoslices[:2], *args, model.split[:3])
    two = hesship.BeckeMolTables(model.R[:2], model.adjust[:2, :2], dev)
    with pytest.raises(hesship.HesshipError, match="full atom list"):
        ni.make_hessian_setup_becke(den, T(D, dev), model.aoslices, T(exc, dev), T(vxc, dev), T(fxc, dev), two,
                                    T(model.grid, dev), T(model.wquad, dev), model.split)
