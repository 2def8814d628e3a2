"""GPU parity: RI-JK skeleton Hessian and first-derivative Fock terms (hh_rijk_skeleton) against
the oracle (the reference's naive term lists, pinned by autograd on the analytic integral model),
RHF and UHF, with several auxiliary ranges, atom lists, J-only / K-only scales and both
handle constructors.  FP64: 1e-10 relative to max|ref|."""
import numpy as np
import pytest

from oracle import rijk as orc
from oracle import rijk_skeleton as osk
from tests.rijk_model import Model
from tests.test_gpu_cphf import T, rel

pytestmark = pytest.mark.gpu


def setup(dev, model):
    I = model.integrals()
    j3c_tp = orc.pack_triu(I["j3c"])
    cderi = orc.cderi_from_j3c(j3c_tp, I["j2c"])
    dev_ints = {k: T(v, dev) for k, v in I.items() if k in ("ip1", "ip2", "ipip1", "ipvip1", "ip1ip2", "ipip2")}
    int3c = lambda kind, p0, p1: dev_ints[kind][:, :, p0:p1]
    return I, j3c_tp, cderi, int3c


def ranges(naux, sizes):
    out, p = [], 0
    for s in sizes:
        if p >= naux:
            break
        out.append((p, min(naux, p + s)))
        p += s
    if p < naux:
        out.append((p, naux))
    return out


@pytest.mark.parametrize("sj,sk", [(1.0, 1.0), (1.0, 0.0), (0.0, 1.0), (0.8, 0.3)])
def test_rhf_skeleton_from_j3c(dev, sj, sk):
    import hesship
    model = Model(nbas=(4, 3, 5), naux_per=(6, 5, 7), seed=2)
    I, j3c_tp, cderi, int3c = setup(dev, model)
    C, occ = model.orbitals(4, seed=3)
    h = hesship.RHessRIJK.from_j3c(T(j3c_tp, dev), T(I["j2c"], dev), sj, sk)
    de, d1 = h.make_skeleton(T(C, dev), occ, int3c, T(I["j2c_ip1"], dev), T(I["j2c_ipip1"], dev),
                             T(I["j2c_ip1ip2"], dev), ranges(model.naux, [5, 7, 4]), model.aoslices, model.auxslices)
    de_o, d1_o = osk.skeleton_rhf(I, C, occ, sj, sk, model.aoslices, model.auxslices)
    assert rel(de.cpu().numpy(), de_o) < 1e-10
    assert rel(d1.cpu().numpy(), d1_o) < 1e-10


def test_rhf_skeleton_from_eig_decomposed_handle(dev):
    """A handle whose cderi came from the eigen decomposition (J^-1/2 kept for the metric terms)
    gives the same skeleton as the Cholesky one on a positive definite metric."""
    import hesship
    model = Model(nbas=(4, 3, 5), naux_per=(6, 5, 7), seed=2)
    I, j3c_tp, cderi, int3c = setup(dev, model)
    C, occ = model.orbitals(4, seed=3)
    h = hesship.RHessRIJK.from_j3c(T(j3c_tp, dev), T(I["j2c"], dev), 1.0, 0.7, threshold=1e-13, policy="eig")
    de, d1 = h.make_skeleton(T(C, dev), occ, int3c, T(I["j2c_ip1"], dev), T(I["j2c_ipip1"], dev),
                             T(I["j2c_ip1ip2"], dev), ranges(model.naux, [5, 7, 4]), model.aoslices, model.auxslices)
    de_o, d1_o = osk.skeleton_rhf(I, C, occ, 1.0, 0.7, model.aoslices, model.auxslices)
    assert rel(de.cpu().numpy(), de_o) < 1e-9
    assert rel(d1.cpu().numpy(), d1_o) < 1e-9


def test_rhf_skeleton_from_cderi_atom_subset(dev):
    """Handle made from cderi (the metric passed to the call); a reordered subset of atoms; one
    auxiliary range."""
    import hesship
    model = Model(nbas=(3, 5, 2, 4), naux_per=(5, 6, 4, 6), seed=4)
    I, j3c_tp, cderi, int3c = setup(dev, model)
    C, occ = model.orbitals(5, seed=5)
    sub = [2, 0, 3]
    aos = [model.aoslices[i] for i in sub]
    auxs = [model.auxslices[i] for i in sub]
    h = hesship.RHessRIJK(T(cderi, dev), 1.0, 0.7)
    de, d1 = h.make_skeleton(T(C, dev), occ, int3c, T(I["j2c_ip1"], dev), T(I["j2c_ipip1"], dev),
                             T(I["j2c_ip1ip2"], dev), [(0, model.naux)], aos, auxs, j2c=T(I["j2c"], dev))
    de_o, d1_o = osk.skeleton_rhf(I, C, occ, 1.0, 0.7, aos, auxs)
    assert rel(de.cpu().numpy(), de_o) < 1e-10
    assert rel(d1.cpu().numpy(), d1_o) < 1e-10


def test_uhf_skeleton(dev):
    import hesship
    model = Model(nbas=(4, 3, 5), naux_per=(6, 5, 7), seed=6)
    I, j3c_tp, cderi, int3c = setup(dev, model)
    Ca, _ = model.orbitals(4, seed=7)
    Cb, _ = model.orbitals(3, seed=8)
    occa = np.where(np.arange(model.nao) < 4, 1.0, 0.0)
    occb = np.where(np.arange(model.nao) < 3, 1.0, 0.0)
    h = hesship.UHessRIJK.from_j3c(T(j3c_tp, dev), T(I["j2c"], dev), 1.0, 0.6)
    de, d1 = h.make_skeleton([T(Ca, dev), T(Cb, dev)], [occa, occb], int3c, T(I["j2c_ip1"], dev),
                             T(I["j2c_ipip1"], dev), T(I["j2c_ip1ip2"], dev), ranges(model.naux, [8, 3]),
                             model.aoslices, model.auxslices)
    de_o, d1_o = osk.skeleton_uhf(I, [Ca, Cb], [occa, occb], 1.0, 0.6, model.aoslices, model.auxslices)
    assert rel(de.cpu().numpy(), de_o) < 1e-10
    for s in range(2):
        assert rel(d1[s].cpu().numpy(), d1_o[s]) < 1e-10


def test_skeleton_with_naux_multiple_of_32(dev):
    """naux = 64 (a multiple of 32): the exchange first-derivative products run on the NT FP64
    kernel; RHF and UHF against the oracle, several auxiliary ranges."""
    import hesship
    model = Model(nbas=(4, 3, 5), naux_per=(22, 18, 24), seed=11)
    assert model.naux == 64
    I, j3c_tp, cderi, int3c = setup(dev, model)
    C, occ = model.orbitals(4, seed=12)
    h = hesship.RHessRIJK.from_j3c(T(j3c_tp, dev), T(I["j2c"], dev), 1.0, 0.8)
    de, d1 = h.make_skeleton(T(C, dev), occ, int3c, T(I["j2c_ip1"], dev), T(I["j2c_ipip1"], dev),
                             T(I["j2c_ip1ip2"], dev), ranges(model.naux, [20, 30]), model.aoslices, model.auxslices)
    de_o, d1_o = osk.skeleton_rhf(I, C, occ, 1.0, 0.8, model.aoslices, model.auxslices)
    assert rel(de.cpu().numpy(), de_o) < 1e-10
    assert rel(d1.cpu().numpy(), d1_o) < 1e-10
    Cb, _ = model.orbitals(3, seed=13)
    occa = np.where(np.arange(model.nao) < 4, 1.0, 0.0)
    occb = np.where(np.arange(model.nao) < 3, 1.0, 0.0)
    hu = hesship.UHessRIJK.from_j3c(T(j3c_tp, dev), T(I["j2c"], dev), 1.0, 0.6)
    de, d1 = hu.make_skeleton([T(C, dev), T(Cb, dev)], [occa, occb], int3c, T(I["j2c_ip1"], dev),
                              T(I["j2c_ipip1"], dev), T(I["j2c_ip1ip2"], dev), ranges(model.naux, [25]),
                              model.aoslices, model.auxslices)
    de_o, d1_o = osk.skeleton_uhf(I, [C, Cb], [occa, occb], 1.0, 0.6, model.aoslices, model.auxslices)
    assert rel(de.cpu().numpy(), de_o) < 1e-10
    for s in range(2):
        assert rel(d1[s].cpu().numpy(), d1_o[s]) < 1e-10


def test_skeleton_hessian_is_translation_invariant(dev):
    """The model integrals are translation invariant, so each row of the skeleton Hessian sums to
    zero over the atoms (the reference's test pattern, tests/hessian_rks_becke/test_b3lyp.rs:64-73)."""
    import hesship
    model = Model(nbas=(5, 4, 6, 3), naux_per=(8, 6, 9, 5), seed=9)
    I, j3c_tp, cderi, int3c = setup(dev, model)
    C, occ = model.orbitals(6, seed=10)
    h = hesship.RHessRIJK.from_j3c(T(j3c_tp, dev), T(I["j2c"], dev), 1.0, 1.0)
    de, _ = h.make_skeleton(T(C, dev), occ, int3c, T(I["j2c_ip1"], dev), T(I["j2c_ipip1"], dev),
                            T(I["j2c_ip1ip2"], dev), ranges(model.naux, [9]), model.aoslices, model.auxslices,
                            with_deriv1=False)
    de = de.cpu().numpy()
    assert np.abs(de.sum(axis=3)).max() < 1e-10 * np.abs(de).max()


def test_skeleton_argument_checks(dev):
    import hesship
    model = Model(seed=11)
    I, j3c_tp, cderi, int3c = setup(dev, model)
    C, occ = model.orbitals(3, seed=12)
    h = hesship.RHessRIJK(T(cderi, dev), 1.0, 1.0)
    args = (T(C, dev), occ, int3c, T(I["j2c_ip1"], dev), T(I["j2c_ipip1"], dev), T(I["j2c_ip1ip2"], dev))
    with pytest.raises(hesship.HesshipError, match="j2c is required"):
        h.make_skeleton(*args, [(0, model.naux)], model.aoslices, model.auxslices)
    with pytest.raises(hesship.HesshipError, match="tile"):
        h.make_skeleton(*args, [(0, 3), (4, model.naux)], model.aoslices, model.auxslices, j2c=T(I["j2c"], dev))
