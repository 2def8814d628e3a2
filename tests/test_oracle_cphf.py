"""Pin the CPHF oracle (oracle/cphf.py) by closed forms on a synthetic RI-JK system."""
import numpy as np

from oracle import cphf, rijk
from tests.test_oracle_rijk import synthetic_ri


def system(nao=8, naux=12, nocc=3, natm=2, seed=0):
    rng = np.random.default_rng(seed)
    _, _, cderi = synthetic_ri(nao, naux, seed=seed + 1)
    cderi *= 0.2
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao); occ[:nocc] = 2.0
    e = np.sort(rng.uniform(-1.0, 1.0, nao)); e[nocc:] += 1.0
    f1 = rng.standard_normal((nao, nocc, 3, natm))
    s1 = rng.standard_normal((nao, nocc, 3, natm)) * 0.1
    rb = lambda b: rijk.rhf_response_bra(cderi, C, occ, b, 1.0, 0.5)
    return C, occ, e, f1, s1, rb


def test_cphf_solution_satisfies_equation():
    C, occ, e, f1, s1, rb = system()
    rhs = cphf.dimless_rhs(f1, s1, occ, e)
    st = {}
    mo1 = cphf.solve_dimless(C, occ, e, [rb], rhs, dict(cphf_tol=1e-10), stats=st)
    res = mo1 + cphf.response_dimless(C, occ, e, [rb], mo1) - rhs
    # the convergence signal is the norm of the new (non-normalised) trial vectors, not the
    # residual itself (krylov_block.rs:130-160), so the residual lands a few orders above tol
    assert np.abs(res).max() < 1e-6
    # occupied block is frozen at -s1/2
    np.testing.assert_allclose(mo1[:3], -0.5 * s1[:3], rtol=0, atol=1e-6)
    # dense solve of the same (vir block) linear system
    nmo, nocc = 8, 3
    n = nmo * nocc
    A = np.zeros((n, n))
    for k in range(n):
        v = np.zeros(n); v[k] = 1.0
        A[:, k] = cphf.response_dimless(C, occ, e, [rb], v.reshape(nmo, nocc, 1)).reshape(-1)
    x = np.linalg.solve(np.eye(n) + A, rhs.reshape(n, -1))
    np.testing.assert_allclose(mo1.reshape(n, -1), x, rtol=1e-5, atol=1e-6)


def test_finalize_and_hessian_by_loops():
    C, occ, e, f1, s1, rb = system(seed=3)
    rhs = cphf.dimless_rhs(f1, s1, occ, e)
    mo1 = cphf.solve_dimless(C, occ, e, [rb], rhs, dict(cphf_tol=1e-10))
    mo1f, mo_e1 = cphf.finalize(C, occ, e, [rb], f1, s1, mo1)
    # finalize leaves a converged solution unchanged (to the solver tolerance)
    np.testing.assert_allclose(mo1f, mo1, atol=1e-8)
    de = cphf.cphf_hess(f1, s1, mo1f, mo_e1, occ, e)
    nocc, natm = 3, 2
    eo = e[:nocc]
    ref = np.zeros((3, 3, natm, natm))
    for A in range(natm):
        for B in range(natm):
            for s in range(3):
                for t in range(3):
                    if B <= A:
                        fa, sa, mb, meb = f1[:, :, t, A], s1[:, :, t, A], mo1f[:, :, s, B], mo_e1[:, :, s, B]
                        v = 4 * np.sum(fa * mb) - 4 * np.sum(sa * mb * eo[None, :]) - 2 * np.sum(sa[:nocc] * meb)
                        ref[s, t, B, A] = v
    for A in range(natm):
        for B in range(A):
            ref[:, :, A, B] = ref[:, :, B, A].T
    np.testing.assert_allclose(de, ref, rtol=1e-12, atol=1e-12)
