"""GPU parity of the resident RHF RI-JK response (the response matrix built once per
prepare_response, one product per call) against the FP64 oracle (hess_r.rs:82-186, 1916-1953):
J-only, K-only and hybrid scales, FP64 (1e-10) and BF16x3 (1e-4), a re-prepared handle, and a CPHF
solve that uses it (rscf.rs:278-300) against the per-call operator's solution."""
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
