"""GPU parity through identities that mathematics fixes, on the library's own results.

The same identities that pin the oracle's factors (tests/test_oracle_pins.py) are checked on the
HIP path through the C ABI, so a factor that differs between the oracle and the library, or that
both got wrong, cannot pass:
- the CPHF Hessian (hh_cphf_* / hh_ucphf_*; rscf.rs:104-399, uscf.rs:78-440) is symmetric, and
  mo_e1 symmetric in (i, j), for first-order matrices that come from symmetric AO matrices;
- the RKS / UKS grid response with a constant LDA kernel (hh_numint_rks/uks_response_bra;
  hess_rks.rs:570-605, hess_uks.rs:390-441) equals the RI-J response of the RI-JK handle built on
  the grid "auxiliary" B[mu nu, g] = sqrt(c w_g) phi_mu(g) phi_nu(g) (hh_rijk_response_bra with
  scale_k = 0; hess_r.rs:1916-1953, hess_u.rs:270-313).
"""
import numpy as np
import pytest
import torch

from oracle import rijk as orc
from tests.test_oracle_rijk import synthetic_ri

pytestmark = pytest.mark.gpu


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def asym(de):
    return np.abs(de - de.transpose(1, 0, 3, 2)).max() / np.abs(de).max()


def mo_e1_asym(m):
    return np.abs(m - m.transpose(1, 0, 2, 3)).max() / np.abs(m).max()


def full_blocks(fwd, rev):
    natm = fwd.shape[3]
    de = np.empty_like(fwd)
    for A in range(natm):
        for B in range(natm):
            de[:, :, B, A] = fwd[:, :, B, A] if B <= A else rev[:, :, natm - 1 - B, natm - 1 - A]
    return de


def sym_first_order(rng, C, nocc, natm):
    nao = C.shape[0]
    F = rng.standard_normal((nao, nao, 3, natm))
    F = F + F.transpose(1, 0, 2, 3)
    S = 0.1 * rng.standard_normal((nao, nao, 3, natm))
    S = S + S.transpose(1, 0, 2, 3)
    f1 = np.einsum("um,uvta,vi->mita", C, F, C[:, :nocc])
    s1 = np.einsum("um,uvta,vi->mita", C, S, C[:, :nocc])
    return f1, s1


TIGHT = dict(cphf_tol=1e-11, cphf_lindep=1e-26, cphf_max_cycle=120)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_rhf_cphf_hessian_symmetric_on_device(dev, precision):
    import hesship
    nao, naux, nocc, natm = 24, 50, 5, 3
    rng = np.random.default_rng(11)
    _, _, cd = synthetic_ri(nao, naux, seed=12)
    cd *= 0.15
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    e = np.sort(rng.uniform(-1.0, 0.5, nao))
    e[nocc:] += 1.0
    f1, s1 = sym_first_order(rng, C, nocc, natm)
    rijk = hesship.RHessRIJK(T(cd, dev), 1.0, 0.5)
    rijk.make_response_preparation(T(C, dev), occ)
    drv = hesship.RHessCPHF(T(C, dev), occ, e, TIGHT)
    drv.add_el(rijk)
    f1d, s1d = T(f1, dev), T(s1, dev)
    rhs = drv.compute_dimless_rhs(f1d, s1d)
    mo1 = drv.solve_dimless_cphf(rhs, precision=precision)
    mo1f, mo_e1 = drv.finalize_cphf(f1d, s1d, mo1)
    fwd = drv.get_cphf_hess(f1d, s1d, mo1f, mo_e1).cpu().numpy()
    r = lambda x: x.flip(-1).contiguous()
    rev = drv.get_cphf_hess(r(f1d), r(s1d), r(mo1f), r(mo_e1)).cpu().numpy()
    a = asym(full_blocks(fwd, rev))
    print(f"RHF CPHF Hessian asymmetry ({precision}) {a:.2e}")
    assert a < 1e-8
    assert mo_e1_asym(mo_e1.cpu().numpy()) < 1e-8


def test_uhf_cphf_hessian_symmetric_on_device(dev):
    import hesship
    nao, naux, natm = 20, 44, 3
    noc = (5, 4)
    rng = np.random.default_rng(21)
    _, _, cd = synthetic_ri(nao, naux, seed=22)
    cd *= 0.15
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in noc]
    e = []
    for no in noc:
        x = np.sort(rng.uniform(-1.0, 0.5, nao))
        x[no:] += 1.0
        e.append(x)
    fs = [sym_first_order(rng, C[s], noc[s], natm) for s in range(2)]
    f1 = [T(fs[s][0], dev) for s in range(2)]
    s1 = [T(fs[s][1], dev) for s in range(2)]
    rijk = hesship.UHessRIJK(T(cd, dev), 1.0, 0.5)
    rijk.make_response_preparation([T(c, dev) for c in C], occ)
    drv = hesship.UHessCPHF([T(c, dev) for c in C], occ, e, TIGHT)
    drv.add_el(rijk)
    rhs = drv.compute_dimless_rhs(f1, s1)
    mo1 = drv.solve_dimless_cphf(rhs)
    mo1f, mo_e1 = drv.finalize_cphf(f1, s1, mo1)
    fwd = drv.get_cphf_hess(f1, s1, mo1f, mo_e1).cpu().numpy()
    r = lambda xs: [x.flip(-1).contiguous() for x in xs]
    rev = drv.get_cphf_hess(r(f1), r(s1), r(mo1f), r(mo_e1)).cpu().numpy()
    a = asym(full_blocks(fwd, rev))
    print(f"UHF CPHF Hessian asymmetry {a:.2e}")
    assert a < 1e-8
    assert max(mo_e1_asym(m.cpu().numpy()) for m in mo_e1) < 1e-8


def lda_grid(rng, ng, nao, c):
    ao = rng.standard_normal((ng, nao, 1))
    w = rng.uniform(0.1, 1.0, ng)
    phi = ao[:, :, 0].T
    B = np.sqrt(c * w)[None, None, :] * phi[:, None, :] * phi[None, :, :]
    return ao, w, orc.pack_triu(B)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 2e-5), ("bf16x3", 1e-4)])
def test_rks_constant_lda_equals_grid_ri_j_on_device(dev, precision, tol):
    import hesship
    ng, nao, nocc, nprop, c = 200, 24, 5, 7, 0.37
    rng = np.random.default_rng(31)
    ao, w, cd = lda_grid(rng, ng, nao, c)
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, nprop))
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=64)
    got = ni.get_rks_response_bra("rho", T(np.full((ng, 1, 1), c), dev), T(bra, dev), T(C[:, :nocc], dev),
                                  precision=precision).cpu().numpy()
    rj = hesship.RHessRIJK(T(cd, dev), 1.0, 0.0)
    rj.make_response_preparation(T(C, dev), occ)
    ref = rj.get_response_bra(T(bra, dev)).cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"RKS constant-LDA vs grid RI-J ({precision}) rel err {err:.2e}")
    assert err < tol


def test_uks_constant_lda_equals_grid_ri_j_on_device(dev):
    import hesship
    ng, nao, nprop, c = 180, 20, 6, 0.29
    noc = (5, 3)
    rng = np.random.default_rng(41)
    ao, w, cd = lda_grid(rng, ng, nao, c)
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in noc]
    bras = [T(rng.standard_normal((nao, no, nprop)), dev) for no in noc]
    ni = hesship.NIMatmul(T(ao, dev), T(w, dev), nbatch=64)
    got = ni.get_uks_response_bra("rho", T(np.full((ng, 1, 2, 1, 2), c), dev), bras,
                                  [T(C[s][:, :noc[s]], dev) for s in range(2)])
    rj = hesship.UHessRIJK(T(cd, dev), 1.0, 0.0)
    rj.make_response_preparation([T(x, dev) for x in C], occ)
    ref = rj.get_response_bra(bras)
    err = max(((g - r).abs().max() / r.abs().max()).item() for g, r in zip(got, ref))
    print(f"UKS constant-LDA vs grid RI-J rel err {err:.2e}")
    assert err < 1e-11
