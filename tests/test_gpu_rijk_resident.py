"""GPU parity of the resident RI-JK response (the response matrix built once per
prepare_response, one product per call) against the FP64 oracle (hess_r.rs:82-186, 1916-1953;
UHF hess_u.rs:270-313): J-only, K-only and hybrid scales, FP64 (1e-10) and BF16x3 (1e-4), a
re-prepared handle, CPHF solves that use it (rscf.rs:278-300, uscf.rs:278-349) against the
oracle's solution, and the resident response accumulated after a grid component."""
import numpy as np
import pytest
import torch

from oracle import cphf as ocp
from oracle import rijk as orc
from tests.test_oracle_rijk import synthetic_ri

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "bf16x3": 1e-4}


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
@pytest.mark.parametrize("nao,naux,nocc,nprop", [(9, 17, 3, 4), (45, 70, 7, 13), (100, 90, 21, 6)])
def test_resident_response_against_oracle(dev, precision, nao, naux, nocc, nprop):
    from hesship import RHessRIJK
    rng = np.random.default_rng(nao)
    _, _, cderi = synthetic_ri(nao, naux, seed=nao + 1)
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, nprop))
    for sj, sk in [(1.0, 1.0), (1.0, 0.0), (0.0, 0.5), (0.3, 0.7)]:
        obj = RHessRIJK(T(cderi, dev), sj, sk)
        obj.set_response_path("resident")
        obj.make_response_preparation(T(C, dev), occ)
        got = obj.get_response_bra(T(bra, dev), precision=precision).cpu().numpy()
        ref = orc.rhf_response_bra(cderi, C, occ, bra, sj, sk)
        err = rel(got, ref)
        print(f"resident {precision} nao={nao} nocc={nocc} scales=({sj},{sk}): rel err {err:.2e}")
        assert err < TOL[precision]


def test_resident_follows_reprepare_and_budget(dev):
    from hesship import RHessRIJK
    nao, naux, nocc, nprop = 30, 50, 6, 5
    rng = np.random.default_rng(3)
    _, _, cderi = synthetic_ri(nao, naux, seed=4)
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, nprop))
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    obj.set_response_path("resident")
    for seed in (1, 2):
        C, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((nao, nao)))
        obj.make_response_preparation(T(C, dev), occ)
        ref = orc.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)
        assert rel(obj.get_response_bra(T(bra, dev), precision="fp64").cpu().numpy(), ref) < 1e-10
        assert rel(obj.get_response_bra(T(bra, dev), precision="bf16x3").cpu().numpy(), ref) < 1e-4
    # a budget below the matrix size falls back to the per-call algorithm (same result)
    obj.set_response_path("resident", budget_bytes=1024)
    assert rel(obj.get_response_bra(T(bra, dev), precision="fp64").cpu().numpy(), ref) < 1e-10


def test_cphf_with_resident_response(dev):
    import hesship
    nao, naux, nocc, natm = 24, 50, 5, 3
    rng = np.random.default_rng(0)
    _, _, cderi = synthetic_ri(nao, naux, seed=1)
    cderi *= 0.15
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    e = np.sort(rng.uniform(-1.0, 0.5, nao))
    e[nocc:] += 1.0
    f1 = rng.standard_normal((nao, nocc, 3, natm))
    s1 = rng.standard_normal((nao, nocc, 3, natm)) * 0.1
    rb = lambda b: orc.rhf_response_bra(cderi, C, occ, b, 1.0, 0.5)
    rhs = ocp.dimless_rhs(f1, s1, occ, e)
    ref = ocp.solve_dimless(C, occ, e, [rb], rhs, dict(cphf_tol=1e-10))
    for precision in ("fp64", "bf16x3"):
        rijk = hesship.RHessRIJK(T(cderi, dev), 1.0, 0.5)
        rijk.set_response_path("resident")
        rijk.make_response_preparation(T(C, dev), occ)
        drv = hesship.RHessCPHF(T(C, dev), occ, e, dict(cphf_tol=1e-10))
        drv.add_el(rijk)
        got = drv.solve_dimless_cphf(T(rhs, dev), precision=precision).cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
@pytest.mark.parametrize("nao,naux,noc,nprop", [(9, 17, (3, 2), 4), (45, 70, (7, 6), 13), (64, 96, (12, 9), 5)])
def test_uhf_resident_response_against_oracle(dev, precision, nao, naux, noc, nprop):
    from hesship import UHessRIJK
    rng = np.random.default_rng(nao + 7)
    _, _, cderi = synthetic_ri(nao, naux, seed=nao + 2)
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in noc]
    bra = [rng.standard_normal((nao, no, nprop)) for no in noc]
    for sj, sk in [(1.0, 1.0), (1.0, 0.0), (0.0, 0.5), (0.3, 0.7)]:
        obj = UHessRIJK(T(cderi, dev), sj, sk)
        obj.set_response_path("resident")
        obj.make_response_preparation([T(c, dev) for c in C], occ)
        got = obj.get_response_bra([T(b, dev) for b in bra], precision=precision)
        ref = orc.uhf_response_bra(cderi, C, occ, bra, sj, sk)
        for sp in range(2):
            err = rel(got[sp].cpu().numpy(), ref[sp])
            print(f"UHF resident {precision} nao={nao} nocc={noc} scales=({sj},{sk}) spin {sp}: rel err {err:.2e}")
            assert err < TOL[precision]


def test_ucphf_with_resident_response_after_grid_component(dev):
    """UKS-hybrid response with the grid component first and the resident UHF RI-JK second (its
    product then accumulates onto the grid response), and the stacked Krylov solve, against the
    oracle."""
    import hesship
    from oracle import numint as oni
    from tests.test_gpu_ucphf import usystem
    from tests.test_gpu_uks_numint import uks_case
    cderi, C, occ, e, f1, s1, rb = usystem(nao=24, naux=50, nocc=(5, 4), natm=2, seed=11, sk=0.2)
    ao, w, _, _, f = uks_case(ng=600, nao=24, nocc=(5, 4), nset=1, seed=12, den="sigma")
    f *= 1e-5
    mocc = [C[0][:, :5], C[1][:, :4]]
    # the reference's stopping test is on trial-vector norms (lindep 1e-24 lets it reach 1e-10);
    # the solutions then agree to ~1e-9 (the true residual is about ten times the test's)
    cfg = dict(cphf_tol=1e-10, cphf_lindep=1e-24)
    for precision, tol in (("fp64", 1e-10), ("bf16x3", 1e-4)):
        drv = hesship.UHessCPHF([T(c, dev) for c in C], occ, e, cfg)
        ni = hesship.NIMatmul(T(ao, dev), T(w, dev))
        drv.add_numint(ni, "sigma", T(f, dev))
        rijk = hesship.UHessRIJK(T(cderi, dev), 1.0, 0.2)
        rijk.set_response_path("resident")
        rijk.make_response_preparation([T(c, dev) for c in C], occ)
        drv.add_el(rijk)
        rn = lambda b: oni.uks_response_bra("sigma", f, ao, w, b, mocc)
        rng = np.random.default_rng(13)
        mo1 = [rng.standard_normal(x.shape) for x in f1]
        got = [x.cpu().numpy() for x in drv.response_mo([T(x, dev) for x in mo1], precision=precision)]
        for sp, r in enumerate(ocp.u_response_mo(C, [rb, rn], mo1)):
            assert rel(got[sp], r) < tol
        rhs = ocp.u_dimless_rhs(f1, s1, occ, e)
        x = [t.cpu().numpy() for t in drv.solve_dimless_cphf([T(r, dev) for r in rhs], precision=precision)]
        xo = ocp.u_solve_dimless(C, occ, e, [rb, rn], rhs, cfg)
        for sp in range(2):
            np.testing.assert_allclose(x[sp], xo[sp], rtol=1e-7, atol=1e-8)
