"""GPU parity of the integral / AO engine (hh_cint_*): AO values against the oracle (all
derivative orders, l <= 4); 2- and 3-centre Coulomb integrals of every kind against the oracle
on shell blocks of low total angular momentum, and at higher angular momentum (f AO, g auxiliary)
against central finite differences of the engine itself, translation identities and index
symmetries; the RI-JK skeleton fed by the engine's integrals against the oracle's skeleton and
translation invariance.  The basis is synthetic (no basis-set files exist here)."""
import os

import numpy as np
import pytest

from oracle import cint as oc
from oracle import rijk as orc
from oracle import rijk_skeleton as osk
from tests import cint_basis as cb
from tests.cint_basis import AO, ATOMS, AUX
from tests.test_gpu_cphf import T, rel

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cint_blocks.npz")


def bases(atoms=ATOMS, cart=False):
    import hesship
    return hesship.CInt.from_basis(atoms, AO, cart), hesship.CInt.from_basis(atoms, AUX, cart)


def shells_of(m):
    return cb.shells_of(m.atm, m.bas, m.env, m.cart)


@pytest.mark.parametrize("deriv", [0, 1, 2, 3])
@pytest.mark.parametrize("cart", [False, True])
def test_eval_gto_matches_oracle(dev, deriv, cart):
    mol, _ = bases(cart=cart)
    rng = np.random.default_rng(deriv)
    pts = rng.normal(0.0, 1.3, (150, 3)) + np.array(ATOMS[rng.integers(0, 3)][1])
    got = mol.eval_gto(T(pts, dev), deriv).cpu().numpy()
    sh = [(s[1], s[2], s[3], s[4]) for s in shells_of(mol)]
    ref = oc.eval_gto(sh, pts, deriv, cart=cart)
    assert got.shape == ref.shape
    assert rel(got, ref) < 1e-12


def test_eval_gto_high_l(dev):
    import hesship
    m = hesship.CInt.from_basis([("X", (0.2, -0.1, 0.3))], {"X": [(4, [[0.8, 1.0]]), (5, [[0.6, 1.0]])]})
    pts = np.random.default_rng(5).normal(0.0, 1.5, (60, 3))
    got = m.eval_gto(T(pts, dev), 2).cpu().numpy()
    ref = oc.eval_gto([(s[1], s[2], s[3], s[4]) for s in shells_of(m)], pts, 2)
    assert rel(got, ref) < 1e-12


def _nine_back(got):
    """[.., 3 (s), 3 (t)] -> [.., 9] with c = 3 t + s (the oracle's order)"""
    return got.transpose(tuple(range(got.ndim - 2)) + (got.ndim - 1, got.ndim - 2)).reshape(got.shape[:-2] + (9,))


@pytest.mark.parametrize("kind", list(cb.KINDS3))
def test_int3c2e_blocks_match_golden(dev, kind):
    """Shell blocks of every kind against the oracle's values (tests/golden/cint_blocks.npz)."""
    mol, aux = bases()
    got = mol.int3c2e(aux, kind).cpu().numpy()
    got = _nine_back(got) if got.ndim == 5 else (got[..., None] if got.ndim == 3 else got)
    g = np.load(GOLDEN)
    n = 0
    while f"3c_{kind}_{n}" in g:
        ref, (a0, b0, c0) = g[f"3c_{kind}_{n}"], g[f"3c_{kind}_{n}_at"]
        blk = got[a0:a0 + ref.shape[0], b0:b0 + ref.shape[1], c0:c0 + ref.shape[2]]
        assert np.abs(blk - ref).max() < 1e-11 * max(np.abs(ref).max(), 1e-3), (kind, n, ref.shape)
        n += 1
    assert n > 5


@pytest.mark.parametrize("kind", list(cb.KINDS2))
def test_int2c2e_blocks_match_golden(dev, kind):
    _, aux = bases()
    got = aux.int2c2e(kind).cpu().numpy()
    got = _nine_back(got) if got.ndim == 4 else (got[..., None] if got.ndim == 2 else got)
    g = np.load(GOLDEN)
    n = 0
    while f"2c_{kind}_{n}" in g:
        ref, (a0, c0) = g[f"2c_{kind}_{n}"], g[f"2c_{kind}_{n}_at"]
        blk = got[a0:a0 + ref.shape[0], c0:c0 + ref.shape[1]]
        assert np.abs(blk - ref).max() < 1e-11 * max(np.abs(ref).max(), 1e-3), (kind, n, ref.shape)
        n += 1
    assert n > 5


def _shifted(atom, axis, h):
    at = [(s, tuple(x)) for s, x in ATOMS]
    x = list(at[atom][1])
    x[axis] += h
    at[atom] = (at[atom][0], tuple(x))
    return bases(at)


def _on(m, A):
    sl = m.aoslice_by_atom()
    mask = np.zeros(m.nao, dtype=bool)
    mask[sl[A, 0]:sl[A, 1]] = True
    return mask


def test_int3c2e_derivatives_by_finite_differences(dev):
    """Every derivative kind at the full angular momenta (f AO, g auxiliary): moving atom A moves
    every function on it, so blocks whose other indices sit elsewhere isolate one derivative:
    d/dA (mu nu|P) = -ip1 (mu on A), -ip2 (P on A); d/dA ip1 = -ipip1 (mu on A), d/dB ip1 = -ipvip1
    (nu on B), d/dC ip1 = -ip1ip2 (P on C), d/dC ip2 = -ipip2 (P on C)."""
    mol, aux = bases()
    I = {k: mol.int3c2e(aux, k).cpu().numpy() for k in ("plain", "ip1", "ip2", "ipip1", "ipvip1", "ip1ip2",
                                                          "ipip2")}
    h = 1e-4

    def fd(atom, axis, kind):
        p, q = _shifted(atom, axis, h), _shifted(atom, axis, -h)
        return (p[0].int3c2e(p[1], kind).cpu().numpy() - q[0].int3c2e(q[1], kind).cpu().numpy()) / (2 * h)

    A, B = 0, 1
    mA, mB, xA, xB = _on(mol, A), _on(mol, B), _on(aux, A), _on(aux, B)
    nA, nB = ~mA, ~mB
    for s in range(3):
        d = fd(A, s, "plain")
        ref = -I["ip1"][..., s]
        sel = np.ix_(mA, nA, ~xA)
        assert np.abs(d[sel] - ref[sel]).max() < 1e-7 * np.abs(ref[sel]).max()
        ref2 = -I["ip2"][..., s]
        sel2 = np.ix_(nA, nA, xA)
        assert np.abs(d[sel2] - ref2[sel2]).max() < 1e-7 * np.abs(ref2[sel2]).max()
        dip1 = fd(A, s, "ip1")  # [.., 3 (t)]
        dip2 = fd(B, s, "ip2")
        dipB = fd(B, s, "ip1")
        for t in range(3):
            r = -I["ipip1"][..., s, t]
            assert np.abs(dip1[..., t][sel] - r[sel]).max() < 1e-7 * np.abs(r[sel]).max()
            r = -I["ipvip1"][..., s, t]
            selv = np.ix_(mA, mB, ~xB)
            assert np.abs(dipB[..., t][selv] - r[selv]).max() < 1e-7 * np.abs(r[selv]).max()
            r = -I["ip1ip2"][..., s, t]
            selc = np.ix_(mA, nB, xB)
            assert np.abs(dipB[..., t][selc] - r[selc]).max() < 1e-7 * np.abs(r[selc]).max()
            r = -I["ipip2"][..., s, t]
            sel3 = np.ix_(nB, nB, xB)
            assert np.abs(dip2[..., t][sel3] - r[sel3]).max() < 1e-7 * np.abs(r[sel3]).max()


def test_int2c2e_derivatives_by_finite_differences(dev):
    _, aux = bases()
    I = {k: aux.int2c2e(k).cpu().numpy() for k in ("plain", "ip1", "ipip1", "ip1ip2")}
    h = 1e-4
    xA, xB = _on(aux, 0), _on(aux, 1)
    for s in range(3):
        p, q = _shifted(0, s, h)[1], _shifted(0, s, -h)[1]
        d = (p.int2c2e("plain").cpu().numpy() - q.int2c2e("plain").cpu().numpy()) / (2 * h)
        sel = np.ix_(xA, ~xA)
        assert np.abs(d[sel] + I["ip1"][..., s][sel]).max() < 1e-7 * np.abs(I["ip1"]).max()
        d1 = (p.int2c2e("ip1").cpu().numpy() - q.int2c2e("ip1").cpu().numpy()) / (2 * h)
        pb, qb = _shifted(1, s, h)[1], _shifted(1, s, -h)[1]
        d2 = (pb.int2c2e("ip1").cpu().numpy() - qb.int2c2e("ip1").cpu().numpy()) / (2 * h)
        for t in range(3):
            assert np.abs(d1[..., t][sel] + I["ipip1"][..., s, t][sel]).max() < 1e-7 * np.abs(I["ipip1"]).max()
            selb = np.ix_(xA, xB)
            assert np.abs(d2[..., t][selb] + I["ip1ip2"][..., s, t][selb]).max() < 1e-7 * np.abs(I["ip1ip2"]).max()


def test_translation_identities_and_symmetries(dev):
    mol, aux = bases()
    I = {k: mol.int3c2e(aux, k).cpu().numpy() for k in ("plain", "ip1", "ip2", "ipip1", "ipvip1", "ip1ip2",
                                                          "ipip2")}
    big = max(np.abs(v).max() for v in I.values())
    assert np.abs(I["plain"] - I["plain"].transpose(1, 0, 2)).max() < 1e-14 * big
    s1 = I["ip1"] + I["ip1"].transpose(1, 0, 2, 3) + I["ip2"]
    assert np.abs(s1).max() < 1e-12 * big
    s2 = I["ipip1"] + I["ipvip1"] + I["ip1ip2"]
    assert np.abs(s2).max() < 1e-12 * big
    s3 = I["ip1ip2"] + I["ip1ip2"].transpose(1, 0, 2, 3, 4) + I["ipip2"]
    assert np.abs(s3).max() < 1e-12 * big
    J = {k: aux.int2c2e(k).cpu().numpy() for k in ("plain", "ip1", "ipip1", "ip1ip2")}
    assert np.abs(J["plain"] - J["plain"].T).max() < 1e-14 * np.abs(J["plain"]).max()
    assert np.linalg.eigvalsh(J["plain"]).min() > 0
    assert np.abs(J["ip1"] + J["ip1"].transpose(1, 0, 2)).max() < 1e-12 * np.abs(J["ip1"]).max()
    assert np.abs(J["ipip1"] + J["ip1ip2"]).max() < 1e-12 * np.abs(J["ipip1"]).max()


def test_aux_ranges_and_generator(dev):
    import hesship
    mol, aux = bases()
    full = mol.int3c2e(aux, "ip1ip2").cpu().numpy()
    for sh0, sh1 in [(0, 3), (3, 5), (5, aux.nbas)]:
        part = mol.int3c2e(aux, "ip1ip2", shls=(sh0, sh1)).cpu().numpy()
        p0, p1 = aux.ao_loc[sh0], aux.ao_loc[sh1]
        assert np.array_equal(part, full[:, :, p0:p1])
    g = hesship.Int3cGenerator(mol, aux)
    assert g.fn and g.ctx


def _skeleton_inputs(dev, mol, aux):
    I = {"j3c": mol.int3c2e(aux).cpu().numpy(), "j2c": aux.int2c2e().cpu().numpy()}
    for k in ("ip1", "ip2", "ipip1", "ipvip1", "ip1ip2", "ipip2"):
        I[k] = mol.int3c2e(aux, k).cpu().numpy()
    for k in ("ip1", "ipip1", "ip1ip2"):
        I["j2c_" + k] = aux.int2c2e(k).cpu().numpy()
    return I


def test_skeleton_on_engine_integrals(dev):
    """The RI-JK skeleton Hessian with the engine's integrals: equal to the oracle's skeleton on the
    same integrals, and translation invariant (each row sums to zero over the atoms)."""
    import hesship
    mol, aux = bases()
    I = _skeleton_inputs(dev, mol, aux)
    nao, naux = mol.nao, aux.nao
    rng = np.random.default_rng(3)
    C = np.linalg.qr(rng.standard_normal((nao, nao)))[0]
    occ = np.where(np.arange(nao) < 5, 2.0, 0.0)
    j3c_tp = orc.pack_triu(I["j3c"])
    h = hesship.RHessRIJK.from_j3c(T(j3c_tp, dev), T(I["j2c"], dev), 1.0, 1.0)

    def int3c(kind, p0, p1):
        sh0 = int(np.nonzero(aux.ao_loc == p0)[0][0])
        sh1 = int(np.nonzero(aux.ao_loc == p1)[0][0])
        return mol.int3c2e(aux, kind, shls=(sh0, sh1))

    ranges = [(int(aux.ao_loc[0]), int(aux.ao_loc[4])), (int(aux.ao_loc[4]), int(aux.ao_loc[-1]))]
    de, d1 = h.make_skeleton(T(C, dev), occ, int3c, T(I["j2c_ip1"], dev), T(I["j2c_ipip1"], dev),
                             T(I["j2c_ip1ip2"], dev), ranges, mol.aoslice_by_atom(), aux.aoslice_by_atom())
    de = de.cpu().numpy()
    de_o, d1_o = osk.skeleton_rhf(I, C, occ, 1.0, 1.0, mol.aoslice_by_atom(), aux.aoslice_by_atom())
    assert rel(de, de_o) < 1e-10
    assert rel(d1.cpu().numpy(), d1_o) < 1e-10
    assert np.abs(de.sum(axis=3)).max() < 1e-9 * np.abs(de).max()


def test_int3c2e_s2ij_packs_the_upper_triangle(dev):
    mol, aux = bases()
    full = mol.int3c2e(aux).cpu().numpy()
    got = mol.int3c2e_s2ij(aux).cpu().numpy()
    assert np.array_equal(got, orc.pack_triu(full))
    part = mol.int3c2e_s2ij(aux, shls=(2, 6)).cpu().numpy()
    assert np.array_equal(part, got[:, aux.ao_loc[2]:aux.ao_loc[6]])
