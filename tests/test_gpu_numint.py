"""GPU parity: numint grid contractions and the xceff transform against the oracle.

FP64 tolerance 1e-11 relative to max|ref|; mixed modes 1e-5 (fp32) / 1e-4 (bf16x3).
"""
import numpy as np
import pytest
import torch

from oracle import cphf as ocp
from oracle import numint as oni
from oracle import rijk as orc
from oracle import xceff as oxc
from tests.test_oracle_numint_xceff import grid_case

pytestmark = pytest.mark.gpu


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
@pytest.mark.parametrize("nbatch", [0, 256])
def test_rho1_and_fxc_bra_trans(dev, den, nbatch):
    import hesship
    ao, w, bra, kets, f = grid_case(ng=700, nao=30, nocc=4, nset=5, seed=3, den=den)
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=nbatch)
    rho1 = ni.make_rho_from_one_bra_mult_ket(T(bra, dev), T(kets, dev), den).cpu().numpy()
    ref = oni.rho1_one_bra_mult_ket(ao, bra, kets, den)
    assert rel(rho1, ref) < 1e-11
    out = ni.make_rks_fxc_pot_with_eff_bra_trans(T(f, dev), T(ref, dev), T(bra, dev), den).cpu().numpy()
    assert rel(out, oni.fxc_pot_bra_trans(den, f, ref, ao, w, bra)) < 1e-11


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
def test_rks_response_bra(dev, den):
    import hesship
    ao, w, bra, kets, f = grid_case(ng=1000, nao=36, nocc=5, nset=6, seed=4, den=den)
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=384)
    kets4 = kets.reshape(36, 5, 3, 2)
    got = ni.get_rks_response_bra(den, T(f, dev), T(kets4, dev), T(bra, dev)).cpu().numpy()
    assert got.shape == kets4.shape
    assert rel(got, oni.rks_response_bra(den, f, ao, w, kets4, bra)) < 1e-11


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_rks_response_mixed(dev, prec, tol):
    import hesship
    ao, w, bra, kets, f = grid_case(ng=1500, nao=40, nocc=6, nset=4, seed=5, den="sigma")
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev))
    got = ni.get_rks_response_bra("sigma", T(f, dev), T(kets, dev), T(bra, dev), precision=prec).cpu().numpy()
    e = rel(got, oni.rks_response_bra("sigma", f, ao, w, kets, bra))
    assert 1e-14 < e < tol, e


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
@pytest.mark.parametrize("nocc,nset,nbatch", [(6, 4, 0), (5, 23, 96), (40, 5, 0), (50, 3, 160), (97, 2, 0),
                                              (130, 2, 0)])
def test_rks_response_bf16x3_fused_matches_unfused(dev, den, nocc, nset, nbatch):
    """The BF16X3 response with the kets product fused into the pointwise pass equals the
    unfused kernels' result bit for bit (tiles of 1..8 perturbations, partial tiles, batches
    that end inside a 32-point block; nocc 130 falls back to the unfused path), and both match
    the oracle at the BF16x3 level."""
    import hesship
    nao = 37
    ao, w, bra, kets, f = grid_case(ng=1111, nao=nao, nocc=nocc, nset=nset, seed=nocc + nset, den=den)
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=nbatch)
    ni.set_fused(True)
    fused = ni.get_rks_response_bra(den, T(f, dev), T(kets, dev), T(bra, dev), precision="bf16x3")
    ni.set_fused(False)
    plain = ni.get_rks_response_bra(den, T(f, dev), T(kets, dev), T(bra, dev), precision="bf16x3")
    assert torch.equal(fused, plain)
    assert rel(fused.cpu().numpy(), oni.rks_response_bra(den, f, ao, w, kets, bra)) < 1e-4


@pytest.mark.parametrize("den", ["sigma", "tau"])
def test_xceff_transform(dev, den):
    import hesship
    rng = np.random.default_rng(6)
    ng = 1001
    nv, xl = oxc.NVAR[den], oxc.XLEN[den]
    rho = rng.standard_normal((ng, nv))
    xc = rng.standard_normal((ng, 1 + xl + xl * (xl + 1) // 2))
    for order in (1, 2):
        got = hesship.transform_xc(T(rho, dev), T(xc, dev), den, order).cpu().numpy()
        np.testing.assert_allclose(got, oxc.transform(rho, xc, den, order), rtol=1e-14, atol=1e-14)


def test_xceff_lda(dev):
    import hesship
    xc = np.random.default_rng(7).standard_normal((50, 3))
    for order in (1, 2):
        got = hesship.transform_xc(None, T(xc, dev), "rho", order).cpu().numpy()
        np.testing.assert_array_equal(got.reshape(-1), xc[:, order])


def test_cphf_response_with_rijk_and_numint(dev):
    """RKS-hybrid response: C^T [RI-JK(scale_k) + 4 fxc(rho1)] C through the CPHF driver."""
    import hesship
    from tests.test_gpu_cphf import system
    cderi, C, occ, e, f1, s1, rb = system(nao=24, naux=50, nocc=5, natm=2, seed=8, sk=0.2)
    ao, w, _, _, f = grid_case(ng=600, nao=24, nocc=5, nset=1, seed=9, den="sigma")
    # scale so the grid response is comparable to the RI-JK one; at 1e-3 the random indefinite
    # kernel makes I + A ill-posed and the oracle's own Krylov solve diverges
    f *= 1e-5
    mocc = C[:, :5]
    rijk = hesship.RHessRIJK(T(cderi, dev), 1.0, 0.2)
    rijk.make_response_preparation(T(C, dev), occ)
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev))
    drv = hesship.RHessCPHF(T(C, dev), occ, e, dict(cphf_tol=1e-10))
    drv.add_el(rijk)
    drv.add_numint(ni, "sigma", T(f, dev))
    mo1 = np.random.default_rng(10).standard_normal(f1.shape)
    rn = lambda b: oni.rks_response_bra("sigma", f, ao, w, b, mocc)
    got = drv.response_mo(T(mo1, dev)).cpu().numpy()
    assert rel(got, ocp.response_mo(C, [rb, rn], mo1)) < 1e-10
    rhs = ocp.dimless_rhs(f1, s1, occ, e)
    x = drv.solve_dimless_cphf(T(rhs, dev)).cpu().numpy()
    xo = ocp.solve_dimless(C, occ, e, [rb, rn], rhs, dict(cphf_tol=1e-10))
    np.testing.assert_allclose(x, xo, rtol=1e-7, atol=1e-9)
