"""GPU parity: dimensionless CPHF driver (rscf.rs) with the RI-JK response, against the oracle."""
import numpy as np
import pytest
import torch

from oracle import cphf as ocp
from oracle import rijk as orc
from tests.test_oracle_rijk import synthetic_ri

pytestmark = pytest.mark.gpu


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def system(nao=24, naux=50, nocc=5, natm=3, seed=0, sk=0.5):
    rng = np.random.default_rng(seed)
    _, _, cderi = synthetic_ri(nao, naux, seed=seed + 1)
    cderi *= 0.15
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao); occ[:nocc] = 2.0
    e = np.sort(rng.uniform(-1.0, 0.5, nao)); e[nocc:] += 1.0
    f1 = rng.standard_normal((nao, nocc, 3, natm))
    s1 = rng.standard_normal((nao, nocc, 3, natm)) * 0.1
    rb = lambda b: orc.rhf_response_bra(cderi, C, occ, b, 1.0, sk)
    return cderi, C, occ, e, f1, s1, rb


def make(dev, cderi, C, occ, e, sk=0.5, cfg=None):
    import hesship
    rijk = hesship.RHessRIJK(T(cderi, dev), 1.0, sk)
    rijk.make_response_preparation(T(C, dev), occ)
    drv = hesship.RHessCPHF(T(C, dev), occ, e, cfg)
    drv.add_el(rijk)
    return drv, rijk


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_rhs_and_responses(dev):
    cderi, C, occ, e, f1, s1, rb = system()
    drv, _ = make(dev, cderi, C, occ, e)
    rhs = drv.compute_dimless_rhs(T(f1, dev), T(s1, dev)).cpu().numpy()
    np.testing.assert_allclose(rhs, ocp.dimless_rhs(f1, s1, occ, e), rtol=1e-14, atol=1e-15)
    mo1 = np.random.default_rng(1).standard_normal(f1.shape)
    r = drv.response_mo(T(mo1, dev)).cpu().numpy()
    assert rel(r, ocp.response_mo(C, [rb], mo1)) < 1e-10
    r = drv.response_dimless_cphf(T(mo1, dev)).cpu().numpy()
    assert rel(r, ocp.response_dimless(C, occ, e, [rb], mo1)) < 1e-10


def test_solve_finalize_hessian_fp64(dev):
    cderi, C, occ, e, f1, s1, rb = system(seed=2)
    drv, _ = make(dev, cderi, C, occ, e, cfg=dict(cphf_tol=1e-10))
    rhs_o = ocp.dimless_rhs(f1, s1, occ, e)
    so = {}
    mo1_o = ocp.solve_dimless(C, occ, e, [rb], rhs_o, dict(cphf_tol=1e-10), stats=so)
    st = {}
    mo1 = drv.solve_dimless_cphf(T(rhs_o, dev), stats=st).cpu().numpy()
    assert st["cycles"] == so["cycles"]
    np.testing.assert_allclose(mo1, mo1_o, rtol=1e-7, atol=1e-9)
    m_o, e1_o = ocp.finalize(C, occ, e, [rb], f1, s1, mo1_o)
    m, e1 = drv.finalize_cphf(T(f1, dev), T(s1, dev), T(mo1, dev))
    np.testing.assert_allclose(m.cpu().numpy(), m_o, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(e1.cpu().numpy(), e1_o, rtol=1e-7, atol=1e-9)
    de_o = ocp.cphf_hess(f1, s1, m_o, e1_o, occ, e)
    de = drv.get_cphf_hess(T(f1, dev), T(s1, dev), m, e1).cpu().numpy()
    np.testing.assert_allclose(de, de_o, rtol=1e-7, atol=1e-8)
    # the same from the one-call driver
    de2 = drv.make_cphf_hess(T(f1, dev), T(s1, dev)).cpu().numpy()
    np.testing.assert_allclose(de2, de_o, rtol=1e-7, atol=1e-8)


def test_cphf_hess_assembly_exact(dev):
    """get_cphf_hess alone, on arbitrary inputs: identical algebra to the oracle's loops."""
    cderi, C, occ, e, f1, s1, rb = system(seed=3)
    drv, _ = make(dev, cderi, C, occ, e)
    rng = np.random.default_rng(4)
    mo1 = rng.standard_normal(f1.shape)
    moe = rng.standard_normal((5, 5, 3, 3))
    de = drv.get_cphf_hess(T(f1, dev), T(s1, dev), T(mo1, dev), T(moe, dev)).cpu().numpy()
    np.testing.assert_allclose(de, ocp.cphf_hess(f1, s1, mo1, moe, occ, e), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_refinement_reaches_fp64_tolerance(dev, prec):
    cderi, C, occ, e, f1, s1, rb = system(seed=5)
    drv, _ = make(dev, cderi, C, occ, e, cfg=dict(cphf_tol=1e-9))
    rhs = ocp.dimless_rhs(f1, s1, occ, e)
    st = {}
    mo1 = drv.solve_dimless_cphf(T(rhs, dev), precision=prec, stats=st).cpu().numpy()
    x_ref = ocp.solve_dimless(C, occ, e, [rb], rhs, dict(cphf_tol=1e-11))
    # FP64 residual of the refined solution
    res = mo1 + ocp.response_dimless(C, occ, e, [rb], mo1) - rhs
    assert np.sqrt((res.reshape(-1, res.shape[2] * res.shape[3]) ** 2).sum(0)).max() < 1e-9 * 1.0001
    np.testing.assert_allclose(mo1, x_ref, rtol=1e-6, atol=1e-8)
    assert st["refinements"] >= 1 and st["residual"] < 1e-9


def test_finalize_reuses_mixed_solve_response(dev):
    """finalize after a mixed-precision solve takes the response of mo1 from the solve's final
    residual (FP64 up to small corrections whose reduced-precision response is within 5 % of
    cphf_tol): the same results, to that level, as evaluating it again in FP64 (after an FP64
    solve the cache is empty) and as the oracle's finalize; a changed mo1 is not reused."""
    cderi, C, occ, e, f1, s1, rb = system(seed=8)
    drv, _ = make(dev, cderi, C, occ, e, cfg=dict(cphf_tol=1e-9))
    rhs = T(ocp.dimless_rhs(f1, s1, occ, e), dev)
    F1, S1 = T(f1, dev), T(s1, dev)
    mo1 = drv.solve_dimless_cphf(rhs, precision="bf16x3")
    m_a, e_a = drv.finalize_cphf(F1, S1, mo1)
    assert drv.finalize_reuses() == 1
    m_o, e_o = ocp.finalize(C, occ, e, [rb], f1, s1, mo1.cpu().numpy())
    np.testing.assert_allclose(m_a.cpu().numpy(), m_o, rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose(e_a.cpu().numpy(), e_o, rtol=1e-8, atol=1e-9)
    bumped = mo1.clone()
    bumped[7, 1, 0, 0] += 1e-12
    drv.finalize_cphf(F1, S1, bumped)
    assert drv.finalize_reuses() == 1
    drv.solve_dimless_cphf(rhs)  # FP64: no cached response
    m_b, e_b = drv.finalize_cphf(F1, S1, mo1)
    assert drv.finalize_reuses() == 1
    np.testing.assert_allclose(m_b.cpu().numpy(), m_a.cpu().numpy(), rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose(e_b.cpu().numpy(), e_a.cpu().numpy(), rtol=1e-8, atol=1e-9)


def test_right_preconditioner_gives_same_solution(dev):
    cderi, C, occ, e, f1, s1, rb = system(seed=6)
    drv, _ = make(dev, cderi, C, occ, e, cfg=dict(cphf_tol=1e-10))
    rhs = T(ocp.dimless_rhs(f1, s1, occ, e), dev)
    x0 = drv.solve_dimless_cphf(rhs).cpu().numpy()
    d = torch.full((24, 5), 0.8, dtype=torch.float64, device=dev)
    x1 = drv.solve_dimless_cphf(rhs, diag=d).cpu().numpy()
    np.testing.assert_allclose(x1, x0, rtol=1e-6, atol=1e-8)


def test_graph_replay_matches_eager(dev):
    """The response operator replayed from a HIP graph: four calls on the same operand addresses
    with new input values written (by torch, eagerly) between them, each compared with the eager
    operator; then whole solves with and without graphs."""
    stream = torch.cuda.Stream(device=dev)    # graphs need a created stream (not the default one)
    with torch.cuda.stream(stream):
        _graph_replay_body(dev)
    torch.cuda.synchronize()
    # on the legacy default stream the operator runs eagerly
    cderi, C_, occ, e, f1, s1, rb = system(seed=5)
    drv, _ = make(dev, cderi, C_, occ, e)
    drv.solve_dimless_cphf(T(ocp.dimless_rhs(f1, s1, occ, e), dev))
    assert drv.graph_stats() == (0, 0)


def _graph_replay_body(dev):
    from hesship import _lib
    from hesship.tensors import empty_f, fortran
    cderi, C_, occ, e, f1, s1, rb = system(seed=5)
    drv, _ = make(dev, cderi, C_, occ, e)
    L = _lib.lib()
    nmo, nocc = C_.shape[1], int((occ > 0).sum())
    nprop = f1.shape[2] * f1.shape[3]
    x = fortran(torch.zeros((nmo, nocc, nprop), dtype=torch.float64, device=dev))
    out = empty_f((nmo, nocc, nprop), dev)
    rng = np.random.default_rng(11)
    inputs = [rng.standard_normal((nmo, nocc, nprop)) for _ in range(4)]

    def call():
        _lib.check(L.hh_cphf_response_dimless(drv._h, x.data_ptr(), nprop, out.data_ptr(), 0))
        torch.cuda.current_stream().synchronize()
        return out.cpu().numpy().copy()

    drv.set_graphs(False)
    ref = []
    for v in inputs:
        x.copy_(T(v, dev))
        ref.append(call())
    drv.set_graphs(True)
    for v, r in zip(inputs, ref):
        x.copy_(T(v, dev))
        x.mul_(2.0).mul_(0.5)          # eager work on the operand between replays
        got = call()
        assert rel(got, r) < 1e-14
    assert drv.graph_stats() == (1, 3)   # eager, capture + launch, two replays
    # whole solves: identical cycles and solutions
    rhs = T(ocp.dimless_rhs(f1, s1, occ, e), dev)
    st0, st1 = {}, {}
    drv.set_graphs(False)
    a = drv.solve_dimless_cphf(rhs, stats=st0).cpu().numpy()
    drv.set_graphs(True)
    c0 = drv.graph_stats()
    b = drv.solve_dimless_cphf(rhs, stats=st1).cpu().numpy()
    c1 = drv.graph_stats()
    assert st0["cycles"] == st1["cycles"] and c1[1] > c0[1]
    assert rel(b, a) < 1e-13


def test_graph_cache_follows_reprepare_and_new_component(dev):
    """A response graph must not be replayed after the component it captured changed: re-preparing
    the RI-JK component with other orbitals (its reduced-precision copies are made on the first
    eager call and skipped by the graph), or adding a component, gives the eager result."""
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        _graph_reprepare_body(dev)
    torch.cuda.synchronize()


def _graph_reprepare_body(dev):
    import hesship
    from hesship import _lib
    from hesship.tensors import empty_f, fortran
    cderi, C_, occ, e, f1, s1, rb = system(seed=6)
    drv, rijk = make(dev, cderi, C_, occ, e)
    L = _lib.lib()
    nmo, nocc = C_.shape[1], int((occ > 0).sum())
    nprop = f1.shape[2] * f1.shape[3]
    rng = np.random.default_rng(12)
    x = fortran(T(rng.standard_normal((nmo, nocc, nprop)), dev))
    out = empty_f((nmo, nocc, nprop), dev)
    BF16X3 = 2

    def call(graphs):
        drv.set_graphs(graphs)
        _lib.check(L.hh_cphf_response_dimless(drv._h, x.data_ptr(), nprop, out.data_ptr(), BF16X3))
        torch.cuda.current_stream().synchronize()
        return out.cpu().numpy().copy()

    for _ in range(3):           # eager, capture + launch, replay
        call(True)
    assert drv.graph_stats()[1] >= 2
    C2, _ = np.linalg.qr(np.random.default_rng(13).standard_normal(C_.shape))
    rijk.make_response_preparation(T(C2, dev), occ)
    got = [call(True) for _ in range(3)]
    ref = call(False)
    assert rel(ref, call(False)) == 0.0
    before = drv.graph_stats()
    for g in got:
        assert rel(g, ref) < 1e-12
    # the re-prepared component really changed the operator
    rijk.make_response_preparation(T(C_, dev), occ)
    assert rel(call(False), ref) > 1e-3
    # a component added after capture is part of the replayed operator
    for _ in range(3):
        call(True)
    extra = hesship.RHessRIJK(T(cderi, dev), 0.5, 0.0)
    extra.make_response_preparation(T(C_, dev), occ)
    drv.add_el(extra)
    got = [call(True) for _ in range(3)]
    ref = call(False)
    for g in got:
        assert rel(g, ref) < 1e-12
    assert drv.graph_stats()[0] > before[0]
