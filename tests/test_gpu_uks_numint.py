"""GPU parity: UKS grid response (spin-coupled kernel), the spin-polarised xceff transform and
the unrestricted CPHF driver with a grid component, against the oracle.

FP64 tolerance 1e-11 relative to max|ref|; mixed modes 1e-5 (fp32) / 1e-4 (bf16x3).
"""
import numpy as np
import pytest

from oracle import cphf as ocp
from oracle import numint as oni
from oracle import xceff as oxc
from tests.test_gpu_cphf import T, rel

pytestmark = pytest.mark.gpu


def uks_case(ng=700, nao=30, nocc=(4, 3), nset=5, seed=0, den="tau"):
    rng = np.random.default_rng(seed)
    ao = rng.standard_normal((ng, nao, 4))
    w = rng.uniform(0.1, 1.0, ng)
    bra = [rng.standard_normal((nao, n)) for n in nocc]
    kets = [rng.standard_normal((nao, n, nset)) for n in nocc]
    nv = oni.NVAR[den]
    f = rng.standard_normal((ng, nv, 2, nv, 2))
    f = f + f.transpose(0, 3, 4, 1, 2)
    return ao, w, bra, kets, f


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
@pytest.mark.parametrize("nbatch", [0, 256])
def test_uks_fxc_bra_trans(dev, den, nbatch):
    import hesship
    ao, w, bra, kets, f = uks_case(den=den, seed=1)
    rng = np.random.default_rng(2)
    rho1 = rng.standard_normal((ao.shape[0], oni.NVAR[den], 2, 5))
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=nbatch)
    got = ni.make_uks_fxc_pot_with_eff_bra_trans(T(f, dev), T(rho1, dev), [T(b, dev) for b in bra], den)
    ref = oni.uks_fxc_pot_bra_trans(den, f, rho1, ao, w, bra)
    for s in range(2):
        assert rel(got[s].cpu().numpy(), ref[s]) < 1e-11


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
def test_uks_response_bra(dev, den):
    import hesship
    ao, w, bra, kets, f = uks_case(ng=1000, nao=36, nocc=(5, 2), nset=6, seed=3, den=den)
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=384)
    k4 = [k.reshape(36, k.shape[1], 3, 2) for k in kets]
    got = ni.get_uks_response_bra(den, T(f, dev), [T(k, dev) for k in k4], [T(b, dev) for b in bra])
    ref = oni.uks_response_bra(den, f, ao, w, k4, bra)
    for s in range(2):
        assert tuple(got[s].shape) == k4[s].shape
        assert rel(got[s].cpu().numpy(), ref[s]) < 1e-11


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_uks_response_mixed(dev, prec, tol):
    import hesship
    ao, w, bra, kets, f = uks_case(ng=1500, nao=40, nocc=(6, 5), nset=4, seed=4, den="sigma")
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev))
    got = ni.get_uks_response_bra("sigma", T(f, dev), [T(k, dev) for k in kets], [T(b, dev) for b in bra],
                                  precision=prec)
    ref = oni.uks_response_bra("sigma", f, ao, w, kets, bra)
    for s in range(2):
        e = rel(got[s].cpu().numpy(), ref[s])
        assert 1e-14 < e < tol, e


def test_uks_closed_shell_equals_rks(dev):
    """Equal spins and a spin-symmetric kernel: each UKS spin block equals half the RKS response
    with the kernel f_aa + f_ab (the closed-shell reduction of hess_uks.rs to hess_rks.rs)."""
    import hesship
    rng = np.random.default_rng(5)
    ng, nao, nocc, nv = 800, 28, 4, 4
    ao = rng.standard_normal((ng, nao, 4))
    w = rng.uniform(0.1, 1.0, ng)
    mocc = rng.standard_normal((nao, nocc))
    mo1 = rng.standard_normal((nao, nocc, 3))
    faa = rng.standard_normal((ng, nv, nv)); faa = faa + faa.transpose(0, 2, 1)
    fab = rng.standard_normal((ng, nv, nv)); fab = fab + fab.transpose(0, 2, 1)
    fu = np.zeros((ng, nv, 2, nv, 2))
    fu[:, :, 0, :, 0] = fu[:, :, 1, :, 1] = faa
    fu[:, :, 0, :, 1] = fu[:, :, 1, :, 0] = fab
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev))
    u = ni.get_uks_response_bra("sigma", T(fu, dev), [T(mo1, dev)] * 2, [T(mocc, dev)] * 2)
    r = ni.get_rks_response_bra("sigma", T((faa + fab) / 2.0, dev), T(mo1, dev), T(mocc, dev)).cpu().numpy()
    for s in range(2):
        assert rel(u[s].cpu().numpy(), r) < 1e-11


@pytest.mark.parametrize("den", ["sigma", "tau"])
def test_xceff_transform_polarized(dev, den):
    import hesship
    rng = np.random.default_rng(6)
    ng = 1001
    nv, xl = oxc.NVAR_POL[den], oxc.XLEN_POL[den]
    rho = rng.standard_normal((ng, nv, 2))
    xc = rng.standard_normal((ng, 1 + xl + xl * (xl + 1) // 2))
    for order in (1, 2):
        got = hesship.transform_xc_polarized(T(rho, dev), T(xc, dev), den, order).cpu().numpy()
        np.testing.assert_allclose(got, oxc.transform_pol(rho, xc, den, order), rtol=1e-13, atol=1e-13)


def test_xceff_polarized_lda(dev):
    import hesship
    xc = np.random.default_rng(7).standard_normal((50, 6))
    rho = np.random.default_rng(8).uniform(0.1, 1.0, (50, 1, 2))
    for order in (1, 2):
        got = hesship.transform_xc_polarized(None, T(xc, dev), "rho", order).cpu().numpy()
        np.testing.assert_array_equal(got, oxc.transform_pol(rho, xc, "rho", order))


def test_ucphf_with_rijk_and_numint(dev):
    """UKS-hybrid response through the unrestricted driver: C_s^T [RI-JK + 2 fxc(rho1)]_s C_s,
    then the stacked Krylov solve."""
    import hesship
    from tests.test_gpu_ucphf import umake, usystem
    cderi, C, occ, e, f1, s1, rb = usystem(nao=24, naux=50, nocc=(5, 4), natm=2, seed=11, sk=0.2)
    ao, w, _, _, f = uks_case(ng=600, nao=24, nocc=(5, 4), nset=1, seed=12, den="sigma")
    f *= 1e-5  # keeps I + A well posed with a random indefinite kernel
    mocc = [C[0][:, :5], C[1][:, :4]]
    cfg = dict(cphf_tol=1e-10)
    drv, _ = umake(dev, cderi, C, occ, e, sk=0.2, cfg=cfg)
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev))
    drv.add_numint(ni, "sigma", T(f, dev))
    rn = lambda b: oni.uks_response_bra("sigma", f, ao, w, b, mocc)
    rng = np.random.default_rng(13)
    mo1 = [rng.standard_normal(x.shape) for x in f1]
    got = [x.cpu().numpy() for x in drv.response_mo([T(x, dev) for x in mo1])]
    for s, r in enumerate(ocp.u_response_mo(C, [rb, rn], mo1)):
        assert rel(got[s], r) < 1e-10
    rhs = ocp.u_dimless_rhs(f1, s1, occ, e)
    x = [t.cpu().numpy() for t in drv.solve_dimless_cphf([T(r, dev) for r in rhs])]
    xo = ocp.u_solve_dimless(C, occ, e, [rb, rn], rhs, cfg)
    for s in range(2):
        np.testing.assert_allclose(x[s], xo[s], rtol=1e-7, atol=1e-9)
