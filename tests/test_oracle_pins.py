"""Pin the oracle's literal factors by identities that mathematics fixes (CPU).

Three identities, each with a planted-mistake check (the factor is changed in the oracle through
monkeypatch and the identity must then fail):

- CPHF Hessian symmetry.  With first-order Fock and overlap matrices that come from symmetric AO
  matrices, the response part of the Hessian is the quadratic form b_A^T H^-1 b_B of the
  (symmetric) orbital Hessian, so the block for (A, B) computed directly, without the reference's
  mirror (rscf.rs:392-397), equals the transpose of the block for (B, A).  This pins the factors
  4 / -4 / -2 of get_cphf_hess (rscf.rs:387-391) and the frozen occupied block -1/2 s1 of the
  right-hand side (rscf.rs:178); the unrestricted twins (uscf.rs:165-174, 409-440) likewise.
- Symmetry of mo_e1.  mo_e1 is the derivative of the occupied-occupied Fock block, so for the same
  inputs mo_e1[i, j] == mo_e1[j, i]: the -s1_ij e_j of b1 and the mo1_ij (e_i - e_j) term of
  finalize_cphf (rscf.rs:350, uscf.rs:369-407) combine to -1/2 s1_ij (e_i + e_j) only with the
  right sign and factor.  (In the Hessian that term drops out: s1_oo is symmetric and e_i - e_j
  antisymmetric.)
- A constant LDA kernel f(g) = c makes the grid response a Coulomb-like response: with the grid
  "auxiliary" B[mu nu, g] = sqrt(c w_g) phi_mu(g) phi_nu(g), the RKS response (hess_rks.rs:603,
  factor 4) equals the RI-J response of the RHF component with scale_j = 1, scale_k = 0
  (hess_r.rs:1916-1953), whose factors are pinned by the naive einsum form (hess_r_naive.rs:
  886-901).
- The same for UKS (hess_uks.rs:436-439, factor 2) against the UHF RI-J response
  (hess_u.rs:270-313, pinned by hess_u_naive.rs:115-167), with the kernel equal in all four spin
  blocks (a functional of the total density).
"""
import numpy as np
import pytest

from oracle import cphf, numint as oni, rijk
from tests.test_oracle_rijk import synthetic_ri

TIGHT = dict(cphf_tol=1e-12, cphf_lindep=1e-28, cphf_max_cycle=200)


def sym_ao_system(nao=8, naux=12, nocc=3, natm=2, seed=5):
    """RHF system with f1mo / s1mo = C^T X_AO C_occ for random symmetric AO matrices X_AO."""
    rng = np.random.default_rng(seed)
    _, _, cd = synthetic_ri(nao, naux, seed=seed + 1)
    cd *= 0.2
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    e = np.sort(rng.uniform(-1.0, 1.0, nao))
    e[nocc:] += 1.0
    F = rng.standard_normal((nao, nao, 3, natm))
    F = F + F.transpose(1, 0, 2, 3)
    S = 0.1 * rng.standard_normal((nao, nao, 3, natm))
    S = S + S.transpose(1, 0, 2, 3)
    f1 = np.einsum("um,uvta,vi->mita", C, F, C[:, :nocc])
    s1 = np.einsum("um,uvta,vi->mita", C, S, C[:, :nocc])
    rb = lambda b: rijk.rhf_response_bra(cd, C, occ, b, 1.0, 0.5)
    return C, occ, e, f1, s1, rb


def full_blocks(hess, args):
    """All (B, A) blocks computed directly: the forward call gives B <= A, the call on the reversed
    atom order gives B >= A (its own B' <= A'); the mirrored copies are never used."""
    fwd = hess(*[a[..., :] for a in args[:4]], *args[4:])
    rev = hess(*[a[..., ::-1] for a in args[:4]], *args[4:])
    natm = fwd.shape[3]
    de = np.empty_like(fwd)
    for A in range(natm):
        for B in range(natm):
            de[:, :, B, A] = fwd[:, :, B, A] if B <= A else rev[:, :, natm - 1 - B, natm - 1 - A]
    return de


def asym(de):
    return np.abs(de - de.transpose(1, 0, 3, 2)).max() / np.abs(de).max()


def rhf_cphf_solved():
    C, occ, e, f1, s1, rb = sym_ao_system()
    rhs = cphf.dimless_rhs(f1, s1, occ, e)
    mo1 = cphf.solve_dimless(C, occ, e, [rb], rhs, TIGHT)
    mo1f, mo_e1 = cphf.finalize(C, occ, e, [rb], f1, s1, mo1)
    return f1, s1, mo1f, mo_e1, occ, e


def rhf_cphf_asym():
    return asym(full_blocks(cphf.cphf_hess, rhf_cphf_solved()))


def mo_e1_asym(mo_e1):
    return np.abs(mo_e1 - mo_e1.transpose(1, 0, 2, 3)).max() / np.abs(mo_e1).max()


def test_rhf_cphf_hessian_is_symmetric():
    assert rhf_cphf_asym() < 1e-10


def test_rhf_mo_e1_is_symmetric():
    assert mo_e1_asym(rhf_cphf_solved()[3]) < 1e-10


def _mo_e1_dropped_eij(b1, mo1, e_ij, nocc):
    return b1[:nocc]


def _mo_e1_flipped_eij(b1, mo1, e_ij, nocc):
    return b1[:nocc] - mo1[:nocc] * e_ij


PLANTED_R = [
    ("HESS_R", (2.0, -4.0, -2.0)),
    ("HESS_R", (4.0, -2.0, -2.0)),
    ("HESS_R", (4.0, -4.0, -1.0)),
    ("HESS_R", (4.0, 4.0, -2.0)),
    ("RHS_OCC", -1.0),
    ("RHS_OCC", 0.0),
]


@pytest.mark.parametrize("name,value", PLANTED_R, ids=[f"{n}-{i}" for i, (n, _) in enumerate(PLANTED_R)])
def test_rhf_cphf_symmetry_catches_planted_mistake(monkeypatch, name, value):
    monkeypatch.setattr(cphf, name, value)
    assert rhf_cphf_asym() > 1e-4


@pytest.mark.parametrize("mutant", [_mo_e1_dropped_eij, _mo_e1_flipped_eij])
def test_mo_e1_symmetry_catches_planted_mistake(monkeypatch, mutant):
    monkeypatch.setattr(cphf, "mo_e1_of", mutant)
    assert mo_e1_asym(rhf_cphf_solved()[3]) > 1e-3


def sym_ao_usystem(nao=8, naux=12, nocc=(3, 2), natm=2, seed=9):
    rng = np.random.default_rng(seed)
    _, _, cd = synthetic_ri(nao, naux, seed=seed + 1)
    cd *= 0.2
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in nocc]
    e = []
    for no in nocc:
        x = np.sort(rng.uniform(-1.0, 1.0, nao))
        x[no:] += 1.0
        e.append(x)
    f1, s1 = [], []
    for s, no in enumerate(nocc):
        F = rng.standard_normal((nao, nao, 3, natm))
        F = F + F.transpose(1, 0, 2, 3)
        S = 0.1 * rng.standard_normal((nao, nao, 3, natm))
        S = S + S.transpose(1, 0, 2, 3)
        f1.append(np.einsum("um,uvta,vi->mita", C[s], F, C[s][:, :no]))
        s1.append(np.einsum("um,uvta,vi->mita", C[s], S, C[s][:, :no]))
    rb = lambda b: rijk.uhf_response_bra(cd, C, occ, b, 1.0, 0.5)
    return C, occ, e, f1, s1, rb


def uhf_cphf_solved():
    C, occ, e, f1, s1, rb = sym_ao_usystem()
    rhs = cphf.u_dimless_rhs(f1, s1, occ, e)
    mo1 = cphf.u_solve_dimless(C, occ, e, [rb], rhs, TIGHT)
    mo1f, mo_e1 = cphf.u_finalize(C, occ, e, [rb], f1, s1, mo1)
    return f1, s1, mo1f, mo_e1, occ, e


def uhf_cphf_asym():
    f1, s1, mo1f, mo_e1, occ, e = uhf_cphf_solved()

    def hess(f1a, s1a, m1a, me1a, f1b, s1b, m1b, me1b, occ, e):
        return cphf.u_cphf_hess([f1a, f1b], [s1a, s1b], [m1a, m1b], [me1a, me1b], occ, e)

    fwd = hess(f1[0], s1[0], mo1f[0], mo_e1[0], f1[1], s1[1], mo1f[1], mo_e1[1], occ, e)
    r = lambda x: x[..., ::-1]
    rev = hess(r(f1[0]), r(s1[0]), r(mo1f[0]), r(mo_e1[0]), r(f1[1]), r(s1[1]), r(mo1f[1]), r(mo_e1[1]), occ, e)
    natm = fwd.shape[3]
    de = np.empty_like(fwd)
    for A in range(natm):
        for B in range(natm):
            de[:, :, B, A] = fwd[:, :, B, A] if B <= A else rev[:, :, natm - 1 - B, natm - 1 - A]
    return asym(de)


def test_uhf_cphf_hessian_is_symmetric():
    assert uhf_cphf_asym() < 1e-10


def test_uhf_mo_e1_is_symmetric():
    mo_e1 = uhf_cphf_solved()[3]
    assert max(mo_e1_asym(m) for m in mo_e1) < 1e-10


def test_uhf_mo_e1_symmetry_catches_planted_mistake(monkeypatch):
    monkeypatch.setattr(cphf, "mo_e1_of", _mo_e1_flipped_eij)
    assert min(mo_e1_asym(m) for m in uhf_cphf_solved()[3]) > 1e-3


PLANTED_U = [
    ("HESS_U", (1.0, -2.0, -1.0)),
    ("HESS_U", (2.0, -1.0, -1.0)),
    ("HESS_U", (2.0, -2.0, -2.0)),
    ("RHS_OCC", -1.0),
]


@pytest.mark.parametrize("name,value", PLANTED_U, ids=[f"{n}-{i}" for i, (n, _) in enumerate(PLANTED_U)])
def test_uhf_cphf_symmetry_catches_planted_mistake(monkeypatch, name, value):
    monkeypatch.setattr(cphf, name, value)
    assert uhf_cphf_asym() > 1e-4


def lda_grid_case(ng=23, nao=7, nset=4, c=0.37, seed=0):
    rng = np.random.default_rng(seed)
    ao = rng.standard_normal((ng, nao, 1))
    w = rng.uniform(0.1, 1.0, ng)
    phi = ao[:, :, 0].T  # [nao, ng]
    B = np.sqrt(c * w)[None, None, :] * phi[:, None, :] * phi[None, :, :]
    return rng, ao, w, rijk.pack_triu(B)


def rks_vs_rij():
    rng, ao, w, cd = lda_grid_case()
    nao, nocc, nset, c = 7, 3, 4, 0.37
    C, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    bra = rng.standard_normal((nao, nocc, nset))
    f = np.full((ao.shape[0], 1, 1), c)
    got = oni.rks_response_bra("rho", f, ao, w, bra, C[:, :nocc])
    ref = rijk.rhf_response_bra(cd, C, occ, bra, 1.0, 0.0)
    return np.abs(got - ref).max() / np.abs(ref).max()


def test_rks_constant_lda_response_equals_grid_ri_j():
    assert rks_vs_rij() < 1e-12


@pytest.mark.parametrize("factor", [2.0, 1.0, 8.0])
def test_rks_identity_catches_planted_factor(monkeypatch, factor):
    monkeypatch.setattr(oni, "RKS_RESPONSE", factor)
    assert rks_vs_rij() > 0.1


def uks_vs_urij():
    rng, ao, w, cd = lda_grid_case(seed=3)
    nao, nset, c = 7, 4, 0.37
    nocc = (3, 2)
    C = [np.linalg.qr(rng.standard_normal((nao, nao)))[0] for _ in range(2)]
    occ = [np.where(np.arange(nao) < no, 1.0, 0.0) for no in nocc]
    bras = [rng.standard_normal((nao, no, nset)) for no in nocc]
    f = np.full((ao.shape[0], 1, 2, 1, 2), c)
    got = oni.uks_response_bra("rho", f, ao, w, bras, [C[s][:, :nocc[s]] for s in range(2)])
    ref = rijk.uhf_response_bra(cd, C, occ, bras, 1.0, 0.0)
    return max(np.abs(g - r).max() / np.abs(r).max() for g, r in zip(got, ref))


def test_uks_constant_lda_response_equals_grid_ri_j():
    assert uks_vs_urij() < 1e-12


@pytest.mark.parametrize("factor", [1.0, 4.0])
def test_uks_identity_catches_planted_factor(monkeypatch, factor):
    monkeypatch.setattr(oni, "UKS_RESPONSE", factor)
    assert uks_vs_urij() > 0.1
