"""GPU parity: RI-JK response through the C ABI against the FP64 oracle (oracle/rijk.py).

Tolerances: FP64 path 1e-10 relative to max|ref| (reference arithmetic, reassociated sums);
FP32 path 1e-5; split-BF16x3 path 1e-4 (both relative to max|ref|).
"""
import numpy as np
import pytest
import torch

from oracle import rijk as orc
from tests.test_oracle_rijk import synthetic_ri

pytestmark = pytest.mark.gpu


def system(nao, naux, nocc, seed=0):
    rng = np.random.default_rng(seed)
    _, _, cderi = synthetic_ri(nao, naux, seed=seed + 100)
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    return rng, cderi, C, occ


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nao,naux,nocc,nprop", [(9, 17, 3, 4), (33, 70, 7, 13), (64, 150, 10, 6)])
def test_rhf_response_fp64(dev, nao, naux, nocc, nprop):
    from hesship import RHessRIJK
    rng, cderi, C, occ = system(nao, naux, nocc)
    bra = rng.standard_normal((nao, nocc, nprop))
    for sj, sk in [(1.0, 1.0), (1.0, 0.0), (0.0, 0.5), (0.3, 0.7)]:
        obj = RHessRIJK(T(cderi, dev), sj, sk)
        obj.make_response_preparation(T(C, dev), occ)
        got = obj.get_response_bra(T(bra, dev)).cpu().numpy()
        ref = orc.rhf_response_bra(cderi, C, occ, bra, sj, sk)
        assert rel(got, ref) < 1e-10, (sj, sk, rel(got, ref))


def test_rhf_response_trailing_shape_and_chunking(dev):
    """bra [nao, nocc, 3, natm]; force several perturbation batches for the K contraction."""
    from hesship import RHessRIJK
    nao, naux, nocc, natm = 20, 41, 4, 3
    rng, cderi, C, occ = system(nao, naux, nocc, seed=1)
    bra = rng.standard_normal((nao, nocc, 3, natm))
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    obj.set_batch(4)   # 9 perturbations -> batches of 4, last one ragged
    obj.set_profiling(True)
    obj.make_response_preparation(T(C, dev), occ)
    ref = orc.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)
    got = obj.get_response_bra(T(bra, dev)).cpu().numpy()
    assert got.shape == bra.shape
    assert rel(got, ref) < 1e-10
    t = obj.timing()
    assert t["k_hot_launches"] == 3 and 0 < t["k_hot_ms"] <= t["k_total_ms"] <= t["response_ms"]


def test_separated_core_matches_oracle(dev):
    from hesship import get_rijk_response_bra_separated
    nao, naux, nocc, nprop = 24, 50, 5, 7
    rng, cderi, C, occ = system(nao, naux, nocc, seed=2)
    bra = rng.standard_normal((nao, nocc, nprop))
    j_ref, k_ref = orc.response_bra_separated(cderi, [C], [occ], [bra], True, True)
    j, k = get_rijk_response_bra_separated(T(cderi, dev), [T(C, dev)], [occ], [T(bra, dev)], True, True)
    assert rel(j.cpu().numpy(), j_ref) < 1e-10
    assert rel(k[0].cpu().numpy(), k_ref[0]) < 1e-10


def test_uhf_response_fp64(dev):
    from hesship import UHessRIJK
    nao, naux = 21, 45
    rng, cderi, Ca, _ = system(nao, naux, 4, seed=3)
    Cb, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occa = np.zeros(nao); occa[:4] = 1.0
    occb = np.zeros(nao); occb[[0, 1, 3]] = 1.0   # non-contiguous occupation
    bra = [rng.standard_normal((nao, 4, 3, 2)), rng.standard_normal((nao, 3, 3, 2))]
    obj = UHessRIJK(T(cderi, dev), 1.0, 0.8)
    obj.make_response_preparation([T(Ca, dev), T(Cb, dev)], [occa, occb])
    got = obj.get_response_bra([T(b, dev) for b in bra])
    ref = orc.uhf_response_bra(cderi, [Ca, Cb], [occa, occb], bra, 1.0, 0.8)
    for s in range(2):
        assert rel(got[s].cpu().numpy(), ref[s]) < 1e-10


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_rhf_response_mixed_precision(dev, prec, tol):
    from hesship import RHessRIJK
    nao, naux, nocc, nprop = 48, 120, 8, 9
    rng, cderi, C, occ = system(nao, naux, nocc, seed=4)
    bra = rng.standard_normal((nao, nocc, nprop))
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    obj.make_response_preparation(T(C, dev), occ)
    got = obj.get_response_bra(T(bra, dev), precision=prec).cpu().numpy()
    ref = orc.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)
    e = rel(got, ref)
    assert 1e-13 < e < tol, e


# BF16x3 operand generation: nocc % 8 != 0 (element loads) and == 0 (vector loads), several
# 64-row groups with a 16-row tail, 2 nocc > 128 (two-column-tile variant), column counts
# naux*nocc that are not multiples of the workgroup's columns, and batches starting at a0 > 0
@pytest.mark.parametrize("nao,naux,nocc,nprop,batch", [
    (37, 90, 5, 7, 0), (70, 64, 24, 5, 0), (90, 40, 70, 3, 0), (45, 33, 16, 9, 4), (130, 50, 13, 4, 3)])
def test_rhf_response_bf16x3_shapes(dev, nao, naux, nocc, nprop, batch):
    from hesship import RHessRIJK
    rng, cderi, C, occ = system(nao, naux, nocc, seed=nao + nocc)
    bra = rng.standard_normal((nao, nocc, nprop))
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    if batch:
        obj.set_batch(batch)
    obj.make_response_preparation(T(C, dev), occ)
    got = obj.get_response_bra(T(bra, dev), precision="bf16x3").cpu().numpy()
    ref = orc.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)
    e = rel(got, ref)
    assert 1e-13 < e < 1e-4, e


def test_zero_bra_gives_zero(dev):
    from hesship import RHessRIJK
    rng, cderi, C, occ = system(12, 20, 3, seed=5)
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    obj.make_response_preparation(T(C, dev), occ)
    got = obj.get_response_bra(torch.zeros(12, 3, 2, dtype=torch.float64, device=dev))
    assert torch.count_nonzero(got).item() == 0


def test_response_before_preparation_is_an_error(dev):
    from hesship import RHessRIJK, HesshipError
    rng, cderi, C, occ = system(6, 8, 2, seed=6)
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    obj.nocc = [2]
    with pytest.raises(HesshipError):
        obj.get_response_bra(torch.zeros(6, 2, 1, dtype=torch.float64, device=dev))


def test_pack_unpack_device(dev):
    from hesship import pack_triu_tilde, unpack_triu
    rng = np.random.default_rng(7)
    dm = rng.standard_normal((11, 11, 3))
    tp = pack_triu_tilde(T(dm, dev)).cpu().numpy()
    np.testing.assert_array_equal(tp, orc.pack_triu_tilde(dm))
    sym = dm + dm.transpose(1, 0, 2)
    full = unpack_triu(T(orc.pack_triu(sym), dev), 11).cpu().numpy()
    np.testing.assert_array_equal(full, sym)


def test_cderi_generated_from_j3c_and_metric(dev):
    """RI metric on the device (Cholesky + TRSM, FP64) + response == naive einsum with j2c^-1."""
    from hesship import RHessRIJK, HesshipError
    rng = np.random.default_rng(11)
    nao, naux, nocc = 14, 130, 3
    j3c = rng.standard_normal((nao, nao, naux)); j3c = 0.5 * (j3c + j3c.transpose(1, 0, 2))
    a = rng.standard_normal((naux, naux)); j2c = a @ a.T + naux * np.eye(naux)
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao); occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, 4))
    obj = RHessRIJK.from_j3c(T(orc.pack_triu(j3c), dev), T(j2c, dev), 1.0, 1.0)
    obj.make_response_preparation(T(C, dev), occ)
    got = obj.get_response_bra(T(bra, dev)).cpu().numpy()
    ref = orc.response_bra_naive(j3c, j2c, C, occ, bra, 1.0, 1.0)
    assert rel(got, ref) < 1e-9
    with pytest.raises(HesshipError):
        RHessRIJK.from_j3c(T(orc.pack_triu(j3c), dev), T(-j2c, dev), 1.0, 1.0, threshold=None)


def _decomp_case(policy):
    rng = np.random.default_rng(21)
    nao, naux, nocc = 12, 70, 3
    j3c = rng.standard_normal((nao, nao, naux)); j3c = 0.5 * (j3c + j3c.transpose(1, 0, 2))
    q, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    if policy == "eig":  # two discarded directions (one below the threshold, one negative)
        ev = np.concatenate([[-1e-3, 1e-9], np.geomspace(1e-2, 50.0, naux - 2)])
    else:  # not positive definite: the Cholesky factorisation fails, the eigenvalue floor applies
        ev = np.concatenate([[-0.5, 2e-4], np.geomspace(1e-1, 20.0, naux - 2)])
    j2c = (q * ev) @ q.T
    j2c = 0.5 * (j2c + j2c.T)
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao); occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, 5))
    return j3c, j2c, C, occ, bra


@pytest.mark.parametrize("policy,thr", [("eig", 1e-6), ("cd", 1e-3)])
def test_cderi_metric_fallback_policies(dev, policy, thr):
    """decomp_j2c_eig and the eigenvalue floor of decomp_j2c_cd on the device (Jacobi eigensolver +
    FP64 products) against the oracle's decompositions: the response of the resulting cderi."""
    from hesship import RHessRIJK
    j3c, j2c, C, occ, bra = _decomp_case(policy)
    tp = orc.pack_triu(j3c)
    obj = RHessRIJK.from_j3c(T(tp, dev), T(j2c, dev), 1.0, 1.0, threshold=thr, policy=policy)
    obj.make_response_preparation(T(C, dev), occ)
    got = obj.get_response_bra(T(bra, dev)).cpu().numpy()
    cderi = orc.cderi_from_j3c_decomp(tp, j2c, policy, thr)
    ref = orc.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)
    assert rel(got, ref) < 1e-9


def test_cderi_eig_on_spd_metric_equals_cholesky(dev):
    """On a positive definite metric both policies give the same response (B B^T = j3c j2c^-1 j3c^T)."""
    from hesship import RHessRIJK, HesshipError
    rng = np.random.default_rng(12)
    nao, naux, nocc = 10, 90, 4
    j3c = rng.standard_normal((nao, nao, naux)); j3c = 0.5 * (j3c + j3c.transpose(1, 0, 2))
    a = rng.standard_normal((naux, naux)); j2c = a @ a.T + naux * np.eye(naux)
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao); occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, 3))
    out = []
    for policy in ("cd", "eig"):
        obj = RHessRIJK.from_j3c(T(orc.pack_triu(j3c), dev), T(j2c, dev), 1.0, 1.0, policy=policy)
        obj.make_response_preparation(T(C, dev), occ)
        out.append(obj.get_response_bra(T(bra, dev)).cpu().numpy())
    ref = orc.response_bra_naive(j3c, j2c, C, occ, bra, 1.0, 1.0)
    assert rel(out[0], ref) < 1e-9 and rel(out[1], ref) < 1e-9
    # no threshold: a metric that is not positive definite is an error (the reference panics)
    with pytest.raises(HesshipError):
        RHessRIJK.from_j3c(T(orc.pack_triu(j3c), dev), T(-j2c, dev), 1.0, 1.0, threshold=None)


def test_bf16x3_response_beyond_128_occupied(dev):
    """nocc above both BF16x3 K algorithms' limits (fused: 80, materialised T tile: 2 nocc <= 256):
    the K product runs in FP32, J in BF16x3; the response stays within the mixed tolerance."""
    from hesship import RHessRIJK
    nao, naux, nocc, nprop = 160, 200, 130, 2
    rng, cderi, C, occ = system(nao, naux, nocc, seed=9)
    bra = rng.standard_normal((nao, nocc, nprop))
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    obj.make_response_preparation(T(C, dev), occ)
    got = obj.get_response_bra(T(bra, dev), precision="bf16x3").cpu().numpy()
    ref = orc.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)
    assert rel(got, ref) < 1e-4
