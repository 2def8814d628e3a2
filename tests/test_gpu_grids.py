"""GPU checks of the molecular grids (hesship.grids: Treutler-Ahlrichs x Lebedev with Becke's
partition on the device) and of the chain grids -> engine AO values -> functional -> DFT
skeleton setup: Gaussian charges integrate to their closed forms, the engine's normalised AO
functions have unit norm on the grid, and the Becke grid-shift Hessian built from all of it
equals the oracle's on the same inputs and is translation invariant."""
import math

import numpy as np
import pytest

from oracle import dft_becke as odb
from tests.test_gpu_cphf import T, rel

pytestmark = pytest.mark.gpu

SVP_LIKE = {"O": [(0, [[2266.2, 0.0053], [340.87, 0.0399], [77.36, 0.1785], [21.48, 0.4643], [6.66, 0.4431]]),
                  (0, [[0.81, 1.0]]), (0, [[0.26, 1.0]]), (1, [[17.72, 0.043], [3.86, 0.229], [1.05, 0.509]]),
                  (1, [[0.28, 1.0]]), (2, [[1.2, 1.0]])],
            "H": [(0, [[13.0, 0.0197], [1.96, 0.138], [0.44, 0.478]]), (0, [[0.12, 1.0]]), (1, [[0.8, 1.0]])]}


def water(n, seed):
    from hesship import synth
    z, R = synth.water_cluster(n, box_ang=4.0 + n, seed=seed)
    sym = {8: "O", 1: "H"}
    return z, R, [(sym[int(c)], tuple(x)) for c, x in zip(z, R)]


def test_grid_integrates_gaussian_charges(dev):
    from hesship.grids import Grids
    z, R, _ = water(2, seed=1)
    g = Grids(z, R, nrad=60, nang=194, device=dev)
    coords, w = g.build()
    assert coords.shape == (len(z) * 60 * 194, 3) and bool((w >= 0).all())
    c, w = coords.cpu().numpy(), w.cpu().numpy()
    for a in (0.6, 2.5):
        dens = sum((a / math.pi) ** 1.5 * np.exp(-a * ((c - RA) ** 2).sum(1)) for RA in R)
        assert (w * dens).sum() == pytest.approx(len(z), rel=1e-6)
    # overlap of two s Gaussians on different atoms against (pi/p)^{3/2} exp(-ab/p |AB|^2)
    a, b = 0.7, 1.3
    f = np.exp(-a * ((c - R[0]) ** 2).sum(1) - b * ((c - R[4]) ** 2).sum(1))
    p = a + b
    ref = (math.pi / p) ** 1.5 * math.exp(-a * b / p * ((R[0] - R[4]) ** 2).sum())
    assert (w * f).sum() == pytest.approx(ref, rel=1e-6)


def test_engine_ao_functions_are_normalised_on_the_grid(dev):
    import hesship
    from hesship.grids import Grids
    z, R, atoms = water(2, seed=2)
    mol = hesship.CInt.from_basis(atoms, SVP_LIKE)
    coords, w = Grids(z, R, nrad=75, nang=302, device=dev).build()
    ao = mol.eval_gto(coords, 0)[:, :, 0]
    S = (ao * w[:, None]).T @ ao
    d = S.diagonal().cpu().numpy()
    np.testing.assert_allclose(d, 1.0, rtol=1e-5)  # the quadrature error of the tight core s functions
    assert float(np.linalg.eigvalsh(S.cpu().numpy()).min()) > 0


@pytest.mark.parametrize("xc,den", [("HYB_GGA_XC_B3LYP", "sigma"), ("HYB_MGGA_XC_TPSS0", "tau")])
def test_becke_setup_on_engine_ao_and_grid(dev, xc, den):
    """B3LYP / TPSS0 grid-shift setup (make_hessian_setup_becke, hess_rks_becke.rs:1106-1266)
    with every input made here: engine AO values to third order on the partitioned grid, the
    device functional at the grid density of a synthetic density matrix."""
    import hesship
    from hesship.grids import Grids, treutler_radii_adjust
    z, R, atoms = water(2, seed=3)
    mol = hesship.CInt.from_basis(atoms, SVP_LIKE)
    g = Grids(z, R, nrad=24, nang=50, device=dev)
    coords, w = g.build()
    ao = mol.eval_gto(coords, 3)
    nao = mol.nao
    rng = np.random.default_rng(4)
    C = np.linalg.qr(rng.standard_normal((nao, nao)))[0][:, :10] * 0.3
    D = 2.0 * C @ C.T
    ni = hesship.NIMatmul(ao, w, nbatch=512)
    rho = ni.setup_rho(den, T(D, dev))
    fn = hesship.LibXCFunctional.from_identifier(xc)
    exc, vxc, fxc = hesship.eval_xc_eff_list([(1.0, fn)], rho)
    aosl = mol.aoslice_by_atom()
    got = ni.make_hessian_setup_becke(den, T(D, dev), aosl, exc, vxc, fxc, g.becke, coords,
                                      g.quadrature_weights, g.atm_quad_split, 3)
    de = got["de_xc_skeleton"].cpu().numpy()
    assert np.abs(de.sum(axis=3)).max() < 1e-9 * np.abs(de).max()
    ref = odb.make_hessian_setup_becke(den, ao.cpu().numpy(), w.cpu().numpy(),
                                       g.quadrature_weights.cpu().numpy(), coords.cpu().numpy(), R,
                                       treutler_radii_adjust(z), 3, D, aosl, g.atm_quad_split,
                                       exc.cpu().numpy(), vxc.cpu().numpy(), fxc.cpu().numpy())
    for k in ("de_xc_skeleton", "vmat_deriv1_grid"):
        assert rel(got[k].cpu().numpy(), ref[k]) < 1e-10, k


def test_tpss0_response_on_engine_ao_and_grid(dev):
    """The RKS grid response of a TPSS0 kernel (get_rks_response_bra, TAU) on engine AO values
    over a Becke grid: the functional's kernels against the oracle's at every point, the response
    against the oracle's contraction of the same kernels; tau >= tau_W holds for the density of a
    positive semidefinite D, which the check asserts."""
    import hesship
    from hesship.grids import Grids
    from oracle import numint as oni
    from oracle import xcfunc as X
    z, R, atoms = water(2, seed=5)
    mol = hesship.CInt.from_basis(atoms, SVP_LIKE)
    coords, w = Grids(z, R, nrad=30, nang=110, device=dev).build()
    ao = mol.eval_gto(coords, 1)
    nao, nocc = mol.nao, 10
    rng = np.random.default_rng(6)
    C = np.linalg.qr(rng.standard_normal((nao, nao)))[0] * 0.4
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    ni = hesship.NIMatmul(ao, w, nbatch=512)
    rho = ni.setup_rho("tau", T(D, dev))
    r = rho.cpu().numpy()
    keep = r[:, 0] > 1e-10
    tw = (r[keep, 1:4] ** 2).sum(1) / (8 * r[keep, 0])
    assert bool((r[keep, 4] >= tw * (1 - 1e-10)).all())
    fn = hesship.LibXCFunctional.from_identifier("HYB_MGGA_XC_TPSS0")
    exc, vxc, fxc = hesship.eval_xc_eff_list([(1.0, fn)], rho)
    ref = X.eval_eff(396, r, 1, "tau")
    for g_, r_ in zip((exc, vxc, fxc), ref):
        g_ = g_.cpu().numpy().reshape(len(r), -1)
        r_ = r_.reshape(len(r), -1)
        sc = np.abs(r_).max(1, keepdims=True)
        assert bool((np.abs(g_ - r_) <= 1e-8 * np.abs(r_) + 1e-10 * sc + 1e-14).all())
    kets = rng.standard_normal((nao, nocc, 3, 2))
    mocc = C[:, :nocc]
    got = ni.get_rks_response_bra("tau", fxc, T(kets, dev), T(mocc, dev)).cpu().numpy()
    want = oni.rks_response_bra("tau", fxc.cpu().numpy(), ao.cpu().numpy()[:, :, :4], w.cpu().numpy(), kets, mocc)
    assert rel(got, want) < 1e-11
