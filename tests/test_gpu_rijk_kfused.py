"""GPU parity of the fused BF16x3 RI-K contraction (the first K term with its operand generated
inside the contraction, the second term as one product over (P, j); hess_r.rs:144-183) against
the FP64 oracle and against the materialised-operand path, over the shapes that exercise its
edges: nocc at each register-class boundary (8, 16, 17, 31, 33, 48, 49, 80), nao not a multiple
of 32 and up to 416 rows (larger nao takes the materialised path), column counts that leave partial tiles and waves, auxiliary counts from
one P per slice up, and UHF spins of different occupation.  Tolerance 1e-4 relative to max|ref|
(the BF16x3 scheme); the two BF16x3 paths agree to the same level."""
import numpy as np
import pytest
import torch

from oracle import rijk as orc
from tests.test_oracle_rijk import synthetic_ri

pytestmark = pytest.mark.gpu


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = [  # nao, naux, nocc, nprop
    (40, 30, 8, 5),
    (37, 9, 16, 3),
    (70, 45, 17, 7),
    (96, 64, 31, 4),
    (100, 50, 33, 6),
    (129, 70, 40, 9),
    (160, 33, 48, 3),
    (200, 40, 49, 5),
    (150, 36, 80, 2),
    (416, 24, 40, 4),
    (420, 24, 40, 4),   # beyond the fused limit for nocc > 16: the materialised path
    (512, 20, 12, 11),  # materialised path
]


@pytest.mark.parametrize("nao,naux,nocc,nprop", CASES)
def test_rhf_k_fused_against_oracle(dev, nao, naux, nocc, nprop):
    from hesship import RHessRIJK
    rng = np.random.default_rng(nao + nocc)
    _, _, cderi = synthetic_ri(nao, naux, seed=nao)
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, nprop))
    ref = orc.rhf_response_bra(cderi, C, occ, bra, 0.0, 1.0)
    obj = RHessRIJK(T(cderi, dev), 0.0, 1.0)
    obj.make_response_preparation(T(C, dev), occ)
    got = obj.get_response_bra(T(bra, dev), precision="bf16x3").cpu().numpy()
    err = rel(got, ref)
    obj.set_k_path("materialized")
    alt = obj.get_response_bra(T(bra, dev), precision="bf16x3").cpu().numpy()
    print(f"K fused nao={nao} naux={naux} nocc={nocc} nprop={nprop}: rel err {err:.2e}, "
          f"vs materialised {rel(got, alt):.2e}")
    assert err < 1e-4
    assert rel(got, alt) < 1e-4


def test_uhf_k_fused_against_oracle(dev):
    from hesship import UHessRIJK
    nao, naux, nprop = 90, 40, 6
    noc = (21, 12)
    rng = np.random.default_rng(5)
    _, _, cderi = synthetic_ri(nao, naux, seed=6)
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in noc]
    bra = [rng.standard_normal((nao, no, nprop)) for no in noc]
    ref = orc.uhf_response_bra(cderi, C, occ, bra, 1.0, 0.7)
    obj = UHessRIJK(T(cderi, dev), 1.0, 0.7)
    obj.make_response_preparation([T(c, dev) for c in C], occ)
    got = obj.get_response_bra([T(b, dev) for b in bra], precision="bf16x3")
    for g, r in zip(got, ref):
        assert rel(g.cpu().numpy(), r) < 1e-4


def test_k_fused_reprepare_and_repeat(dev):
    """Calls after a new prepare_response use the new orbitals (the O and X' packs are per
    prepare); repeated calls give bit-identical results (fixed reduction order)."""
    from hesship import RHessRIJK
    nao, naux, nocc, nprop = 64, 30, 20, 5
    rng = np.random.default_rng(9)
    _, _, cderi = synthetic_ri(nao, naux, seed=10)
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, nprop))
    obj = RHessRIJK(T(cderi, dev), 1.0, 1.0)
    for seed in (1, 2):
        C, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((nao, nao)))
        obj.make_response_preparation(T(C, dev), occ)
        a = obj.get_response_bra(T(bra, dev), precision="bf16x3").cpu().numpy()
        b = obj.get_response_bra(T(bra, dev), precision="bf16x3").cpu().numpy()
        assert np.array_equal(a, b)
        assert rel(a, orc.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)) < 1e-4
