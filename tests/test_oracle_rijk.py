"""Pin the RI-JK oracle (oracle/rijk.py) by mathematics, independent of its own formulas.

The reference's own check (tests/hessian_uhf/test_hessian_uhf.rs:119-187) is that the optimized
separated response equals the naive einsum response built from full j3c and an explicit j2c
inverse, J-only / K-only / full, to (rtol 1e-6, atol 1e-8).  The same check runs here on a
synthetic molecule-like system (the reference's integral inputs are not available).
"""
import numpy as np
import pytest

from oracle import rijk


def synthetic_ri(nao, naux, seed=7):
    """j3c [nao, nao, naux] symmetric in (mu, nu); j2c SPD [naux, naux]; cderi = j3c L^-T."""
    rng = np.random.default_rng(seed)
    j3c = rng.standard_normal((nao, nao, naux))
    j3c = 0.5 * (j3c + j3c.transpose(1, 0, 2))
    a = rng.standard_normal((naux, naux))
    j2c = a @ a.T + naux * np.eye(naux)
    cderi = rijk.cderi_from_j3c(rijk.pack_triu(j3c), j2c)
    return j3c, j2c, cderi


def mo_system(nao, nocc, rng):
    q, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    return q, occ


def test_pack_triu_tilde_rule():
    rng = np.random.default_rng(0)
    nao = 6
    dm = rng.standard_normal((nao, nao, 3))
    tp = rijk.pack_triu_tilde(dm)
    for nu in range(nao):
        for mu in range(nu + 1):
            k = nu * (nu + 1) // 2 + mu
            w = 1.0 if mu == nu else 2.0
            np.testing.assert_allclose(tp[k], w * dm[mu, nu], rtol=0, atol=0)
    # trace identity: sum_tp cderi_tp * tilde(D)_tp == sum_{mu nu} B_{mu nu} D_{mu nu} for symmetric D
    s = dm + dm.transpose(1, 0, 2)
    b = rng.standard_normal((nao, nao))
    b = b + b.T
    lhs = rijk.pack_triu(b).T @ rijk.pack_triu_tilde(s)
    rhs = np.einsum("uv,uva->a", b, s)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-13, atol=1e-12)


def test_unpack_roundtrip():
    rng = np.random.default_rng(1)
    m = rng.standard_normal((7, 7, 2))
    m = m + m.transpose(1, 0, 2)
    np.testing.assert_array_equal(rijk.unpack_triu(rijk.pack_triu(m), 7), m)


def test_cderi_metric():
    j3c, j2c, cderi = synthetic_ri(6, 11)
    lhs = cderi @ cderi.T
    j3tp = rijk.pack_triu(j3c)
    rhs = j3tp @ np.linalg.solve(j2c, j3tp.T)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nbatch", [72, 5])
def test_rhf_optimized_equals_naive(nbatch):
    nao, naux, nocc, nprop = 9, 17, 3, 4
    rng = np.random.default_rng(2)
    j3c, j2c, cderi = synthetic_ri(nao, naux)
    C, occ = mo_system(nao, nocc, rng)
    bra = rng.standard_normal((nao, nocc, 2, 2))
    for sj, sk in [(1.0, 0.0), (0.0, 1.0), (1.0, 1.0), (0.7, 0.3)]:
        opt = rijk.rhf_response_bra(cderi, C, occ, bra, sj, sk, nbatch)
        ref = rijk.response_bra_naive(j3c, j2c, C, occ, bra, sj, sk)
        np.testing.assert_allclose(opt, ref, rtol=1e-10, atol=1e-10)
    assert nprop == 4


def test_uhf_optimized_equals_naive():
    nao, naux = 8, 13
    rng = np.random.default_rng(3)
    j3c, j2c, cderi = synthetic_ri(nao, naux, seed=11)
    Ca, _ = mo_system(nao, 3, rng)
    Cb, _ = mo_system(nao, 2, rng)
    occa = np.zeros(nao); occa[:3] = 1.0
    occb = np.zeros(nao); occb[:2] = 1.0
    # the reference test's bra: arange * 0.01, shape [nao, nocc_s, 3, natm]
    mk = lambda no: (np.arange(nao * no * 3 * 2, dtype=float).reshape(2, 3, no, nao).T * 0.01)
    bra = [mk(3), mk(2)]
    for sj, sk in [(1.0, 0.0), (0.0, 1.0), (1.0, 1.0)]:
        opt = rijk.uhf_response_bra(cderi, [Ca, Cb], [occa, occb], bra, sj, sk)
        ref = rijk.uhf_response_bra_naive(j3c, j2c, [Ca, Cb], [occa, occb], bra, sj, sk)
        for s in range(2):
            np.testing.assert_allclose(opt[s], ref[s], rtol=1e-10, atol=1e-10)


def test_uhf_equal_spins_reduces_to_rhf():
    """alpha == beta: UHF J (x0.5 per spin, summed over two spins) equals RHF J; K equal."""
    nao, naux, nocc = 7, 10, 2
    rng = np.random.default_rng(4)
    _, _, cderi = synthetic_ri(nao, naux, seed=5)
    C, occ = mo_system(nao, nocc, rng)
    occ1 = np.where(occ > 0, 1.0, 0.0)
    bra = rng.standard_normal((nao, nocc, 3))
    r = rijk.rhf_response_bra(cderi, C, occ, bra, 1.0, 1.0)
    u = rijk.uhf_response_bra(cderi, [C, C], [occ1, occ1], [bra, bra], 1.0, 1.0)
    np.testing.assert_allclose(u[0], r, rtol=1e-12, atol=1e-12)


def test_response_linear_and_zero():
    nao, naux, nocc = 6, 9, 2
    rng = np.random.default_rng(6)
    _, _, cderi = synthetic_ri(nao, naux, seed=9)
    C, occ = mo_system(nao, nocc, rng)
    b1 = rng.standard_normal((nao, nocc, 2))
    b2 = rng.standard_normal((nao, nocc, 2))
    f = lambda b: rijk.rhf_response_bra(cderi, C, occ, b, 1.0, 0.5)
    np.testing.assert_allclose(f(2.0 * b1 - 3.0 * b2), 2.0 * f(b1) - 3.0 * f(b2), rtol=1e-11, atol=1e-11)
    np.testing.assert_array_equal(f(np.zeros_like(b1)), 0.0)


def test_j_depends_only_on_symmetric_dm1():
    """J_ao from bra equals J_ao from any bra' with the same bra C^T + C bra^T."""
    nao, naux, nocc = 6, 8, 2
    rng = np.random.default_rng(8)
    _, _, cderi = synthetic_ri(nao, naux, seed=10)
    C, occ = mo_system(nao, nocc, rng)
    mocc = C[:, :nocc]
    bra = rng.standard_normal((nao, nocc, 1))
    # adding mocc @ S with S antisymmetric leaves bra C^T + C bra^T unchanged
    S = rng.standard_normal((nocc, nocc)); S = S - S.T
    bra2 = bra + (mocc @ S)[:, :, None]
    j1, _ = rijk.response_bra_separated(cderi, [C], [occ], [bra], True, False)
    j2, _ = rijk.response_bra_separated(cderi, [C], [occ], [bra2], True, False)
    np.testing.assert_allclose(j1, j2, rtol=1e-11, atol=1e-11)


def metric_with_spectrum(naux, evals, seed):
    """j2c = Q diag(evals) Q^T with a random orthogonal Q: the decompositions are pinned by the
    known eigenpairs rather than by a second eigensolver."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    return (q * evals) @ q.T, q


def test_decomp_eig_is_inverse_sqrt_on_kept_eigenvalues():
    naux = 40
    ev = np.concatenate([[-1e-3, 1e-9, 3e-9], np.geomspace(1e-2, 50.0, naux - 3)])
    j2c, q = metric_with_spectrum(naux, ev, 1)
    thr = 1e-6
    w = rijk.decomp_j2c_eig(j2c, thr)
    keep = ev >= thr
    ref = (q[:, keep] * ev[keep] ** -0.5) @ q[:, keep].T
    np.testing.assert_allclose(w, ref, atol=1e-9 * np.abs(ref).max())
    np.testing.assert_allclose(w, w.T, atol=1e-12 * np.abs(w).max())
    # w j2c w is the projector onto the kept eigenvectors
    np.testing.assert_allclose(w @ j2c @ w, q[:, keep] @ q[:, keep].T, atol=1e-9)


def test_decomp_eig_on_spd_metric_matches_cholesky_cderi():
    j3c, j2c, cderi = synthetic_ri(9, 50, seed=3)
    tp = rijk.pack_triu(j3c)
    b_eig = rijk.cderi_from_j3c_decomp(tp, j2c, "eig")
    # B B^T = j3c j2c^-1 j3c^T for either factor (pure_decompose.rs:382-407 pattern)
    np.testing.assert_allclose(b_eig @ b_eig.T, cderi @ cderi.T, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(b_eig @ b_eig.T, tp @ np.linalg.solve(j2c, tp.T), rtol=1e-10, atol=1e-12)


def test_decomp_cd_floor_and_no_threshold():
    naux = 30
    ev = np.concatenate([[-0.5, 2e-4], np.geomspace(1e-1, 20.0, naux - 2)])
    j2c, q = metric_with_spectrum(naux, ev, 2)
    with pytest.raises(np.linalg.LinAlgError):
        rijk.decomp_j2c_cd(j2c, None)
    thr = 1e-3
    L = rijk.decomp_j2c_cd(j2c, thr)
    np.testing.assert_allclose(np.tril(L), L)
    np.testing.assert_allclose(L @ L.T, (q * np.maximum(ev, thr)) @ q.T, atol=1e-12 * 20)
    # a well-conditioned metric is factorised directly (the floor changes nothing)
    _, spd, _ = synthetic_ri(4, 25, seed=5)
    np.testing.assert_allclose(rijk.decomp_j2c_cd(spd, 1e-14), np.linalg.cholesky(spd))
    # a positive definite metric whose factor has a diagonal element <= threshold takes the floor
    ev2 = np.geomspace(1e-6, 10.0, naux)
    j2c2, q2 = metric_with_spectrum(naux, ev2, 4)
    L2 = rijk.decomp_j2c_cd(j2c2, 1e-2)
    np.testing.assert_allclose(L2 @ L2.T, (q2 * np.maximum(ev2, 1e-2)) @ q2.T, atol=1e-11)
