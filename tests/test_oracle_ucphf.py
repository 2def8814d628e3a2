"""Pin the unrestricted CPHF oracle (oracle/cphf.py u_*): dense solve of the stacked two-spin
system, reduction to the restricted driver for a closed shell, Hessian assembly by loops."""
import numpy as np

from oracle import cphf, rijk
from tests.test_oracle_rijk import synthetic_ri


def usystem(nao=8, naux=12, nocc=(3, 2), natm=2, seed=0):
    rng = np.random.default_rng(seed)
    _, _, cderi = synthetic_ri(nao, naux, seed=seed + 1)
    cderi *= 0.2
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in nocc]
    e = []
    for no in nocc:
        x = np.sort(rng.uniform(-1.0, 1.0, nao)); x[no:] += 1.0
        e.append(x)
    f1 = [rng.standard_normal((nao, no, 3, natm)) for no in nocc]
    s1 = [rng.standard_normal((nao, no, 3, natm)) * 0.1 for no in nocc]
    rb = lambda b: rijk.uhf_response_bra(cderi, C, occ, b, 1.0, 0.5)
    return C, occ, e, f1, s1, rb, cderi


def test_ucphf_solution_satisfies_stacked_equation():
    C, occ, e, f1, s1, rb, _ = usystem()
    rhs = cphf.u_dimless_rhs(f1, s1, occ, e)
    mo1 = cphf.u_solve_dimless(C, occ, e, [rb], rhs, dict(cphf_tol=1e-10))
    resp = cphf.u_response_dimless(C, occ, e, [rb], mo1)
    for s, no in enumerate((3, 2)):
        assert np.abs(mo1[s] + resp[s] - rhs[s]).max() < 1e-6
        np.testing.assert_allclose(mo1[s][:no], -0.5 * s1[s][:no], rtol=0, atol=1e-6)
    # dense solve over the stacked [8*3 + 8*2] space: the spins couple through J
    na, nb = 8 * 3, 8 * 2
    n = na + nb
    A = np.zeros((n, n))
    for k in range(n):
        v = np.zeros(n); v[k] = 1.0
        r = cphf.u_response_dimless(C, occ, e, [rb], [v[:na].reshape(8, 3, 1), v[na:].reshape(8, 2, 1)])
        A[:, k] = np.concatenate([r[0].ravel(), r[1].ravel()])
    assert np.abs(A[:na, na:]).max() > 1e-3  # alpha-beta coupling present
    b = np.concatenate([rhs[0].reshape(na, -1), rhs[1].reshape(nb, -1)])
    x = np.linalg.solve(np.eye(n) + A, b)
    np.testing.assert_allclose(np.concatenate([mo1[0].reshape(na, -1), mo1[1].reshape(nb, -1)]), x,
                               rtol=1e-5, atol=1e-6)


def test_closed_shell_ucphf_equals_rcphf():
    nao, nocc, natm = 8, 3, 2
    rng = np.random.default_rng(7)
    _, _, cderi = synthetic_ri(nao, 12, seed=8)
    cderi *= 0.2
    C = np.linalg.qr(rng.standard_normal((nao, nao)))[0]
    e = np.sort(rng.uniform(-1.0, 1.0, nao)); e[nocc:] += 1.0
    f1 = rng.standard_normal((nao, nocc, 3, natm))
    s1 = rng.standard_normal((nao, nocc, 3, natm)) * 0.1
    occ2 = np.where(np.arange(nao) < nocc, 2.0, 0.0)
    occ1 = occ2 / 2
    rb_r = lambda b: rijk.rhf_response_bra(cderi, C, occ2, b, 1.0, 0.5)
    rb_u = lambda b: rijk.uhf_response_bra(cderi, [C, C], [occ1, occ1], b, 1.0, 0.5)
    cfg = dict(cphf_tol=1e-11)
    rhs = cphf.dimless_rhs(f1, s1, occ2, e)
    mo1 = cphf.solve_dimless(C, occ2, e, [rb_r], rhs, cfg)
    mo1f, mo_e1 = cphf.finalize(C, occ2, e, [rb_r], f1, s1, mo1)
    de_r = cphf.cphf_hess(f1, s1, mo1f, mo_e1, occ2, e)
    urhs = cphf.u_dimless_rhs([f1, f1], [s1, s1], [occ1, occ1], [e, e])
    umo1 = cphf.u_solve_dimless([C, C], [occ1, occ1], [e, e], [rb_u], urhs, cfg)
    umo1f, umo_e1 = cphf.u_finalize([C, C], [occ1, occ1], [e, e], [rb_u], [f1, f1], [s1, s1], umo1)
    de_u = cphf.u_cphf_hess([f1, f1], [s1, s1], umo1f, umo_e1, [occ1, occ1], [e, e])
    for s in range(2):
        np.testing.assert_allclose(umo1f[s], mo1f, rtol=1e-7, atol=1e-8)
    np.testing.assert_allclose(de_u, de_r, rtol=1e-7, atol=1e-8)


def test_ucphf_hessian_by_loops():
    C, occ, e, f1, s1, rb, _ = usystem(seed=3)
    rhs = cphf.u_dimless_rhs(f1, s1, occ, e)
    mo1 = cphf.u_solve_dimless(C, occ, e, [rb], rhs, dict(cphf_tol=1e-10))
    mo1f, mo_e1 = cphf.u_finalize(C, occ, e, [rb], f1, s1, mo1)
    de = cphf.u_cphf_hess(f1, s1, mo1f, mo_e1, occ, e)
    natm = 2
    ref = np.zeros((3, 3, natm, natm))
    for A in range(natm):
        for B in range(A + 1):
            for s in range(3):
                for t in range(3):
                    v = 0.0
                    for sp, no in enumerate((3, 2)):
                        eo = e[sp][:no]
                        fa, sa = f1[sp][:, :, t, A], s1[sp][:, :, t, A]
                        mb, meb = mo1f[sp][:, :, s, B], mo_e1[sp][:, :, s, B]
                        v += 2 * np.sum(fa * mb) - 2 * np.sum(sa * mb * eo[None, :]) - np.sum(sa[:no] * meb)
                    ref[s, t, B, A] = v
        for B in range(A):
            ref[:, :, A, B] = ref[:, :, B, A].T
    np.testing.assert_allclose(de, ref, rtol=1e-12, atol=1e-12)
