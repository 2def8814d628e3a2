"""GPU parity: unrestricted dimensionless CPHF driver (uscf.rs) with the UHF RI-JK response,
against the oracle; closed shells through both drivers give the same Hessian."""
import numpy as np
import pytest

from oracle import cphf as ocp
from oracle import rijk as orc
from tests.test_gpu_cphf import T, rel
from tests.test_oracle_rijk import synthetic_ri

pytestmark = pytest.mark.gpu


def usystem(nao=24, naux=50, nocc=(5, 4), natm=3, seed=0, sk=0.5):
    rng = np.random.default_rng(seed)
    _, _, cderi = synthetic_ri(nao, naux, seed=seed + 1)
    cderi *= 0.15
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in nocc]
    e = []
    for no in nocc:
        x = np.sort(rng.uniform(-1.0, 0.5, nao)); x[no:] += 1.0
        e.append(x)
    f1 = [rng.standard_normal((nao, no, 3, natm)) for no in nocc]
    s1 = [rng.standard_normal((nao, no, 3, natm)) * 0.1 for no in nocc]
    rb = lambda b: orc.uhf_response_bra(cderi, C, occ, b, 1.0, sk)
    return cderi, C, occ, e, f1, s1, rb


def umake(dev, cderi, C, occ, e, sk=0.5, cfg=None):
    import hesship
    rijk = hesship.UHessRIJK(T(cderi, dev), 1.0, sk)
    rijk.make_response_preparation([T(c, dev) for c in C], occ)
    drv = hesship.UHessCPHF([T(c, dev) for c in C], occ, e, cfg)
    drv.add_el(rijk)
    return drv, rijk


def np_(pair):
    return [p.cpu().numpy() for p in pair]


def test_urhs_and_responses(dev):
    cderi, C, occ, e, f1, s1, rb = usystem()
    drv, _ = umake(dev, cderi, C, occ, e)
    rhs = np_(drv.compute_dimless_rhs([T(x, dev) for x in f1], [T(x, dev) for x in s1]))
    for s, r in enumerate(ocp.u_dimless_rhs(f1, s1, occ, e)):
        np.testing.assert_allclose(rhs[s], r, rtol=1e-14, atol=1e-15)
    rng = np.random.default_rng(1)
    mo1 = [rng.standard_normal(x.shape) for x in f1]
    got = np_(drv.response_mo([T(x, dev) for x in mo1]))
    for s, r in enumerate(ocp.u_response_mo(C, [rb], mo1)):
        assert rel(got[s], r) < 1e-10
    got = np_(drv.response_dimless_cphf([T(x, dev) for x in mo1]))
    for s, r in enumerate(ocp.u_response_dimless(C, occ, e, [rb], mo1)):
        assert rel(got[s], r) < 1e-10


def test_usolve_finalize_hessian_fp64(dev):
    cderi, C, occ, e, f1, s1, rb = usystem(seed=2)
    cfg = dict(cphf_tol=1e-10)
    drv, _ = umake(dev, cderi, C, occ, e, cfg=cfg)
    rhs_o = ocp.u_dimless_rhs(f1, s1, occ, e)
    mo1_o = ocp.u_solve_dimless(C, occ, e, [rb], rhs_o, cfg)
    st = {}
    mo1 = np_(drv.solve_dimless_cphf([T(x, dev) for x in rhs_o], stats=st))
    assert st["cycles"] > 0
    for s in range(2):
        np.testing.assert_allclose(mo1[s], mo1_o[s], rtol=1e-7, atol=1e-9)
    m_o, e1_o = ocp.u_finalize(C, occ, e, [rb], f1, s1, mo1_o)
    m, e1 = drv.finalize_cphf([T(x, dev) for x in f1], [T(x, dev) for x in s1], [T(x, dev) for x in mo1])
    for s in range(2):
        np.testing.assert_allclose(m[s].cpu().numpy(), m_o[s], rtol=1e-7, atol=1e-9)
        np.testing.assert_allclose(e1[s].cpu().numpy(), e1_o[s], rtol=1e-7, atol=1e-9)
    de_o = ocp.u_cphf_hess(f1, s1, m_o, e1_o, occ, e)
    de = drv.get_cphf_hess([T(x, dev) for x in f1], [T(x, dev) for x in s1], m, e1).cpu().numpy()
    assert rel(de, de_o) < 1e-8


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_usolve_mixed_precision_refined(dev, prec):
    cderi, C, occ, e, f1, s1, rb = usystem(seed=4)
    tol = 1e-9
    drv, _ = umake(dev, cderi, C, occ, e, cfg=dict(cphf_tol=tol))
    rhs = [T(x, dev) for x in ocp.u_dimless_rhs(f1, s1, occ, e)]
    ref = np_(drv.solve_dimless_cphf(rhs))
    st = {}
    got = np_(drv.solve_dimless_cphf(rhs, precision=prec, stats=st))
    assert st["residual"] < tol and st["refinements"] >= 1
    for s in range(2):
        assert np.abs(got[s] - ref[s]).max() < 1e-7


def test_closed_shell_through_both_drivers(dev):
    import hesship
    nao, nocc, natm = 24, 5, 2
    rng = np.random.default_rng(9)
    _, _, cderi = synthetic_ri(nao, 50, seed=10)
    cderi *= 0.15
    C = np.linalg.qr(rng.standard_normal((nao, nao)))[0]
    e = np.sort(rng.uniform(-1.0, 0.5, nao)); e[nocc:] += 1.0
    f1 = rng.standard_normal((nao, nocc, 3, natm))
    s1 = rng.standard_normal((nao, nocc, 3, natm)) * 0.1
    occ2 = np.where(np.arange(nao) < nocc, 2.0, 0.0)
    occ1 = occ2 / 2
    cfg = dict(cphf_tol=1e-11)
    rr = hesship.RHessRIJK(T(cderi, dev), 1.0, 0.5)
    rr.make_response_preparation(T(C, dev), occ2)
    rd = hesship.RHessCPHF(T(C, dev), occ2, e, cfg)
    rd.add_el(rr)
    de_r = rd.make_cphf_hess(T(f1, dev), T(s1, dev)).cpu().numpy()
    ud, _ = umake(dev, cderi, [C, C], [occ1, occ1], [e, e], cfg=cfg)
    de_u = ud.make_cphf_hess([T(f1, dev)] * 2, [T(s1, dev)] * 2).cpu().numpy()
    assert rel(de_u, de_r) < 1e-8


def test_spin_mismatch_is_an_error(dev):
    import hesship
    from hesship._lib import HesshipError
    cderi, C, occ, e, f1, s1, rb = usystem()
    r_rijk = hesship.RHessRIJK(T(cderi, dev), 1.0, 0.5)
    r_rijk.make_response_preparation(T(C[0], dev), np.where(occ[0] > 0, 2.0, 0.0))
    ud = hesship.UHessCPHF([T(c, dev) for c in C], occ, e)
    with pytest.raises(HesshipError):
        ud.add_el(r_rijk)  # an RHF component on the UHF driver


def _doublet_cation_c6(dev):
    """The alkane C6 doublet cation of the batched workload (molecule 21: nao 154, naux 532, 25 / 24
    occupied, 60 perturbations; seeded synthetic RI tensors) on host and device."""
    import torch
    from hesship import synth
    n, i = 6, 21
    nao, naux, nprop = 24 * n + 10, 84 * n + 28, 3 * (3 * n + 2)
    cd = synth.cderi(nao, naux, seed=synth.SEED_B + 16 * i)
    C = synth.mo_coeff(nao, seed=synth.SEED_C + 16 * i)
    occ = [np.where(np.arange(nao) < 4 * n + 1, 1.0, 0.0), np.where(np.arange(nao) < 4 * n, 1.0, 0.0)]
    noc = [int(o.sum()) for o in occ]
    z = synth.normal(0x5EED0005, 0, 2 * nao + sum(2 * nao * no * nprop for no in noc))
    en, f1, s1, p = [], [], [], 2 * nao
    for sp, no in enumerate(noc):
        zz = z[sp * nao:(sp + 1) * nao]
        en.append(np.sort(np.concatenate([-1.4 + 1.1 * np.abs(np.tanh(zz[:no])), 0.2 + 2.5 * np.abs(np.tanh(zz[no:]))])))
        f1.append(z[p:p + nao * no * nprop].reshape(nprop, no, nao).T.copy())
        p += nao * no * nprop
        s1.append(0.1 * z[p:p + nao * no * nprop].reshape(nprop, no, nao).T.copy())
        p += nao * no * nprop
    return dict(nao=nao, noc=noc, cd=cd, C=C, occ=occ, en=en, f1=f1, s1=s1,
                d_cd=torch.from_numpy(cd.T.copy()).to(dev).T, d_C=torch.from_numpy(C).to(dev))


def _udriver(sys_, scale2):
    import hesship
    rijk = hesship.UHessRIJK(sys_["d_cd"] * scale2 ** 0.5, 1.0, 1.0)
    rijk.make_response_preparation([sys_["d_C"], sys_["d_C"]], sys_["occ"])
    drv = hesship.UHessCPHF([sys_["d_C"], sys_["d_C"]], sys_["occ"], sys_["en"], dict(cphf_tol=1e-8, cphf_max_space=42))
    drv.add_el(rijk)
    return drv, rijk


def _lowest_eig(drv, nao, noc, dev):
    """Lowest eigenvalue of the dimensionless response A on the two vir-occ blocks (ARPACK, one
    FP64 response per matvec; A = D^-1 K with K symmetric, so the spectrum is real)."""
    import torch
    from scipy.sparse.linalg import LinearOperator, eigs
    sizes = [(nao - no) * no for no in noc]

    def mv(v):
        xs, p = [], 0
        for no, sz in zip(noc, sizes):
            x = np.zeros((nao, no, 1))
            x[no:, :, 0] = np.real(v[p:p + sz]).reshape(nao - no, no)
            xs.append(torch.from_numpy(x).to(dev))
            p += sz
        ys = drv.response_dimless_cphf(xs)
        return np.concatenate([y[no:, :, 0].cpu().numpy().ravel() for y, no in zip(ys, noc)])

    n = sum(sizes)
    ev = eigs(LinearOperator((n, n), matvec=mv, dtype=np.float64), k=2, which="SR", return_eigenvectors=False,
              tol=1e-6, ncv=24, v0=np.random.default_rng(7).standard_normal(n))
    return float(np.real(ev).min())


def _true_residual(drv, rhs, x):
    """max over columns of ||rhs - (x + A x)||_2, A the FP64 dimensionless response."""
    ax = drv.response_dimless_cphf(x, precision="fp64")
    cols = sum(((r - xx - a) ** 2).sum(dim=(0, 1)) for r, xx, a in zip(rhs, x, ax))
    return cols.max().sqrt().item()


def _ratio_scale(drv, sys_, dev):
    import torch
    rng = np.random.default_rng(1)
    v = [torch.from_numpy(rng.standard_normal((sys_["nao"], no, 4))).to(dev) for no in sys_["noc"]]
    r = drv.response_dimless_cphf(v)
    return (sum(x.norm() ** 2 for x in r) ** 0.5 / sum(x.norm() ** 2 for x in v) ** 0.5).item()


def test_mixed_solve_on_indefinite_system_reports_true_residual(dev):
    """The doublet cation with the RI scale that gives |A v| = 0.3 |v| for random v: A then has
    eigenvalues below -1, so I + A is indefinite (an unstable reference state, which no CPHF
    system of a converged SCF has), and the mixed refinement stalls (a sweep gains less than 10x).
    The solve must then end in the plain FP64 Krylov solve (or reach the tolerance by refinement),
    and the residual it reports must be the true FP64 residual of the returned solution,
    recomputed here from an independent FP64 response."""
    import torch
    sys_ = _doublet_cation_c6(dev)
    drv, _ = _udriver(sys_, 1.0)
    scale2 = 0.3 / _ratio_scale(drv, sys_, dev)
    drv, _ = _udriver(sys_, scale2)
    lo = _lowest_eig(drv, sys_["nao"], sys_["noc"], dev)
    print(f"lowest eigenvalue of A: {lo:.4f}")
    assert lo < -1.0
    rhs = drv.compute_dimless_rhs([torch.from_numpy(x).to(dev) for x in sys_["f1"]],
                                  [torch.from_numpy(x).to(dev) for x in sys_["s1"]])
    st = {}
    got = drv.solve_dimless_cphf(rhs, precision="bf16x3", stats=st)
    res = _true_residual(drv, rhs, got)
    print(f"mixed solve: {st['refinements']} sweeps, fp64_fallback {st['fp64_fallback']}, reported residual "
          f"{st['residual']:.3e}, recomputed {res:.3e}")
    assert st["fp64_fallback"] == 1 or res < 1e-8
    assert abs(res - st["residual"]) <= 1e-6 * res + 1e-14
    assert res < 1e-6


def test_mixed_solve_on_stable_doublet_cation(dev):
    """The same doublet cation with the RI scale shrunk until the lowest eigenvalue of A is -0.75
    (I + A positive definite with smallest eigenvalue 0.25): the mixed solve converges by
    refinement alone (no FP64 fallback) to the FP64 residual tolerance, checked by an independent
    FP64 residual, and agrees with the oracle's CPU solve (uscf.rs:278-349, FP64 block Krylov to
    1e-10) on six perturbations."""
    import torch
    sys_ = _doublet_cation_c6(dev)
    drv, _ = _udriver(sys_, 1.0)
    scale2 = 0.3 / _ratio_scale(drv, sys_, dev)
    drv, _ = _udriver(sys_, scale2)
    lo = _lowest_eig(drv, sys_["nao"], sys_["noc"], dev)
    scale2 *= 0.75 / -lo
    drv, _ = _udriver(sys_, scale2)
    rhs = drv.compute_dimless_rhs([torch.from_numpy(x).to(dev) for x in sys_["f1"]],
                                  [torch.from_numpy(x).to(dev) for x in sys_["s1"]])
    st = {}
    got = drv.solve_dimless_cphf(rhs, precision="bf16x3", stats=st)
    res = _true_residual(drv, rhs, got)
    print(f"mixed solve: {st['refinements']} sweeps, fp64_fallback {st['fp64_fallback']}, residual {res:.3e}")
    assert st["fp64_fallback"] == 0
    assert res < 1e-8
    cols = [0, 17, 59]
    cd = sys_["cd"] * scale2 ** 0.5
    rb = lambda b: orc.uhf_response_bra(cd, [sys_["C"], sys_["C"]], sys_["occ"], b, 1.0, 1.0)
    ref = ocp.u_solve_dimless([sys_["C"], sys_["C"]], sys_["occ"], sys_["en"], [rb],
                              [r[:, :, cols].cpu().numpy() for r in rhs], {"cphf_tol": 1e-10, "cphf_lindep": 1e-24})
    err = max(np.abs(g[:, :, cols].cpu().numpy() - r).max() for g, r in zip(got, ref))
    print(f"max |mixed - oracle| over perturbations {cols}: {err:.2e}")
    assert err < 1e-7
