"""GPU parity: many CPHF systems of uneven size (restricted and unrestricted) solved in one
lockstep batch, each against the oracle's krylov_block solve of the same system (rscf.rs:278-300,
uscf.rs:278-349) and against the single-system driver."""
import numpy as np
import pytest

from oracle import cphf as ocp
from tests.test_gpu_cphf import T, make, rel, system
from tests.test_gpu_ucphf import np_, umake, usystem

pytestmark = pytest.mark.gpu

TOL = 1e-10
CFG = dict(cphf_tol=TOL, cphf_max_cycle=42, cphf_max_space=42)


def _systems(dev):
    """(kind, driver, rijk, rhs, oracle mo1, oracle stats) for uneven RHF / UHF systems."""
    out = []
    specs = [("r", dict(nao=24, naux=50, nocc=5, natm=3, seed=2)),
             ("u", dict(nao=30, naux=64, nocc=(6, 5), natm=2, seed=4)),
             ("r", dict(nao=41, naux=90, nocc=9, natm=4, seed=6)),
             ("u", dict(nao=17, naux=40, nocc=(3, 2), natm=1, seed=8)),
             ("r", dict(nao=64, naux=130, nocc=12, natm=5, seed=10))]
    for kind, sp in specs:
        if kind == "r":
            cderi, C, occ, e, f1, s1, rb = system(**sp)
            drv, rijk = make(dev, cderi, C, occ, e, cfg=CFG)
            rhs = ocp.dimless_rhs(f1, s1, occ, e)
            so = {}
            mo1 = ocp.solve_dimless(C, occ, e, [rb], rhs, CFG, stats=so)
            out.append(("r", drv, rijk, rhs, mo1, so))
        else:
            cderi, C, occ, e, f1, s1, rb = usystem(**sp)
            drv, rijk = umake(dev, cderi, C, occ, e, cfg=CFG)
            rhs = ocp.u_dimless_rhs(f1, s1, occ, e)
            so = {}
            mo1 = ocp.u_solve_dimless(C, occ, e, [rb], rhs, CFG, stats=so)
            out.append(("u", drv, rijk, rhs, mo1, so))
    return out


def test_batch_matches_oracle_and_driver(dev):
    import hesship
    syss = _systems(dev)
    batch = hesship.CPHFBatch(tol=TOL, max_cycle=42, lindep=1e-14, device=dev)
    for kind, drv, rijk, rhs, _, _ in syss:
        batch.add(drv, rijk, T(rhs, dev) if kind == "r" else [T(x, dev) for x in rhs])
    stats = []
    got = batch.solve(stats=stats)
    nsys, cycles, launches = batch.info()
    assert nsys == len(syss)
    assert cycles == max(s["cycles"] for s in stats)
    assert 0 < launches <= 24
    for (kind, drv, rijk, rhs, mo1_o, so), g, st in zip(syss, got, stats):
        g = g.cpu().numpy() if kind == "r" else np_(g)
        ref = [mo1_o] if kind == "r" else mo1_o
        gg = [g] if kind == "r" else g
        for a, b in zip(gg, ref):
            assert rel(a, b) < 1e-7, (kind, rel(a, b))
        # the same algorithm: the cycle count matches the oracle's krylov_block (no restarts)
        assert abs(st["cycles"] - so["cycles"]) <= 1, (st, so)
        assert st["max_innerprod"] < TOL ** 2
        # and the single-system driver on the same device operator
        one = drv.solve_dimless_cphf(T(rhs, dev) if kind == "r" else [T(x, dev) for x in rhs])
        one = [one.cpu().numpy()] if kind == "r" else np_(one)
        for a, b in zip(gg, one):
            assert rel(a, b) < 1e-8


def test_batch_zero_rhs_and_repeat(dev):
    import hesship
    cderi, C, occ, e, f1, s1, rb = system(nao=20, naux=44, nocc=4, natm=2, seed=12)
    drv, rijk = make(dev, cderi, C, occ, e, cfg=CFG)
    rhs = ocp.dimless_rhs(f1, s1, occ, e)
    batch = hesship.CPHFBatch(tol=TOL, max_cycle=42, device=dev)
    batch.add(drv, rijk, T(rhs, dev))
    batch.add(drv, rijk, T(np.zeros_like(rhs), dev))
    st1 = []
    a = [x.cpu().numpy() for x in batch.solve(stats=st1)]
    assert np.all(a[1] == 0.0) and st1[1]["cycles"] == 0
    st2 = []
    b = [x.cpu().numpy() for x in batch.solve(stats=st2)]
    # a second solve of the same batch starts from scratch and is reproducible
    assert np.array_equal(a[0], b[0]) and st1[0]["cycles"] == st2[0]["cycles"]
    mo1_o = ocp.solve_dimless(C, occ, e, [rb], rhs, CFG)
    assert rel(a[0], mo1_o) < 1e-7


def test_batch_not_converged_and_limits(dev):
    import hesship
    from hesship._lib import HesshipError, KrylovNotConverged
    cderi, C, occ, e, f1, s1, rb = system(nao=24, naux=50, nocc=5, natm=3, seed=14)
    drv, rijk = make(dev, cderi, C, occ, e, cfg=CFG)
    rhs = ocp.dimless_rhs(f1, s1, occ, e)
    batch = hesship.CPHFBatch(tol=1e-12, max_cycle=2, device=dev)
    batch.add(drv, rijk, T(rhs, dev))
    stats = []
    with pytest.raises(KrylovNotConverged):
        batch.solve(stats=stats)
    assert stats[0]["cycles"] == 2 and stats[0]["max_innerprod"] >= 1e-24
    # nprop above the one-group limit is rejected
    big = np.zeros((24, 5, 129))
    with pytest.raises(HesshipError):
        batch.add(drv, rijk, T(big, dev))
