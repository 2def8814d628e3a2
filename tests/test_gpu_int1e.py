"""GPU parity of the one-electron integrals (hh_cint_int1e: overlap, kinetic, nuclear attraction
and rinv at a nucleus, with the ip / ipip / ipvip derivatives the Hessian's core and overlap
parts take) against the oracle's shell blocks (oracle/cint.py block1e, pinned by closed forms and
quadrature), and the first derivatives against central finite differences of the engine's own
matrices under nuclear displacements (the basis functions and the nuclear charges move).  The
basis is the engine tests' synthetic one on N, H, H (no basis-set files exist here)."""
import numpy as np
import pytest

from oracle import cint as oc
from tests import cint_basis as cb

pytestmark = pytest.mark.gpu

ATOMS = [("N", (0.0, 0.0, 0.0)), ("H", (0.0, 1.43, 1.11)), ("H", (0.31, -1.37, 0.98))]
BASIS = {"N": cb.AO["X"], "H": cb.AO["H"]}
MAXL = {"plain": 4, "ip1": 3, "ipip1": 2, "ipvip1": 2}


def mol(atoms=ATOMS):
    import hesship
    return hesship.CInt.from_basis(atoms, BASIS)


def pairs(m, maxl, nmax=14):
    sh = cb.shells_of(m.atm, m.bas, m.env)
    allp = [(a, b) for a in sh for b in sh if a[2] + b[2] <= maxl]
    step = max(1, -(-len(allp) // nmax))
    return allp[::step]


def _flat(t):
    t = t.cpu().numpy()
    return t.reshape(t.shape[0], t.shape[1], -1, order="F") if t.ndim > 2 else t[:, :, None]


@pytest.mark.parametrize("kind", ["plain", "ip1", "ipip1", "ipvip1"])
def test_int1e_blocks_match_oracle(dev, kind):
    m = mol()
    sfx = {"plain": ("int1e_{}",), "ip1": ("int1e_ip{}",), "ipip1": ("int1e_ipip{}",), "ipvip1": ("int1e_ip{}ip",)}[kind]
    Z = m.atom_charges()
    R = m.atom_coords()
    got = {op: _flat(m.intor(sfx[0].format(op))) for op in ("ovlp", "kin", "nuc")}
    got_r = {C: _flat(m.intor(sfx[0].format("rinv"), rinv_atom=C)) for C in range(len(R))}
    # errors relative to the largest element of the same matrix (blocks may vanish by symmetry)
    err = {}

    def note(key, g, r, full):
        err[key] = max(err.get(key, 0.0), float(np.abs(g - r).max() / np.abs(full).max()))
    for a, b in pairs(m, MAXL[kind]):
        sa, sb = (a[2], a[3], a[4]), (b[2], b[3], b[4])
        na, nb = 2 * a[2] + 1, 2 * b[2] + 1
        sl = (slice(a[5], a[5] + na), slice(b[5], b[5] + nb))
        for op in ("ovlp", "kin"):
            note(op, got[op][sl], oc.block1e(op, [a[1], b[1]], [sa, sb], kind), got[op])
        vsum = 0.0
        for C in range(len(R)):
            ref = oc.block1e("rinv", [a[1], b[1]], [sa, sb], kind, C=R[C])
            note(f"rinv{C}", got_r[C][sl], ref, got_r[C])
            vsum = vsum - Z[C] * ref
        note("nuc", got["nuc"][sl], vsum, got["nuc"])
    assert max(err.values()) < 1e-12, err


def test_int1e_first_derivatives_by_finite_differences(dev):
    """d/dR_A of S, T and V: -(ip rows on A) - (ip rows on A)^T, plus for V the moving charge
    -Z_A (iprinv_A + iprinv_A^T)."""
    m = mol()
    nao = m.nao
    aosl = m.aoslice_by_atom()
    ip = {op: m.intor(f"int1e_ip{op}").cpu().numpy() for op in ("ovlp", "kin", "nuc")}
    Z = m.atom_charges()
    h = 1e-4
    for A in range(len(ATOMS)):
        p0, p1 = aosl[A]
        rinv = m.intor("int1e_iprinv", rinv_atom=A).cpu().numpy()
        for x in range(3):
            vals = {}
            for sgn in (1, -1):
                at = [(s, tuple(np.array(c) + (sgn * h if (i == A) else 0.0) * np.eye(3)[x])) for i, (s, c) in
                      enumerate(ATOMS)]
                mm = mol(at)
                vals[sgn] = {op: mm.intor(f"int1e_{op}").cpu().numpy() for op in ("ovlp", "kin", "nuc")}
            for op in ("ovlp", "kin", "nuc"):
                fd = (vals[1][op] - vals[-1][op]) / (2 * h)
                an = np.zeros((nao, nao))
                an[p0:p1] -= ip[op][p0:p1, :, x]
                an[:, p0:p1] -= ip[op][p0:p1, :, x].T
                if op == "nuc":
                    an -= Z[A] * (rinv[:, :, x] + rinv[:, :, x].T)
                assert np.abs(fd - an).max() < 1e-6 * max(1.0, np.abs(an).max()), (A, x, op)


def test_int1e_symmetries(dev):
    m = mol()
    for op in ("ovlp", "kin", "nuc"):
        s = m.intor(f"int1e_{op}").cpu().numpy()
        assert np.abs(s - s.T).max() < 1e-13 * np.abs(s).max()
        ipip = _flat(m.intor(f"int1e_ipip{op}"))
        # d^2 / dt ds on one function commutes: components 3t + s and 3s + t agree
        for t in range(3):
            for u in range(3):
                assert np.abs(ipip[:, :, 3 * t + u] - ipip[:, :, 3 * u + t]).max() < 1e-12 * np.abs(ipip).max()
        ipvip = _flat(m.intor(f"int1e_ip{op}ip"))
        # <d_t a| O |d_s b> = <d_s b| O |d_t a>: the transpose swaps the directions
        for t in range(3):
            for u in range(3):
                assert np.abs(ipvip[:, :, 3 * t + u] - ipvip[:, :, 3 * u + t].T).max() < 1e-12 * np.abs(ipvip).max()
    S = m.intor("int1e_ovlp").cpu().numpy()
    assert np.abs(np.diag(S) - 1.0).max() < 1e-12 and np.linalg.eigvalsh(S).min() > 0
