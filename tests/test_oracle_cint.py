"""Pins of the integral / AO oracle (oracle/cint.py) by mathematics that does not go through it:
spherical-harmonic orthonormality and harmonicity, libcint's tabulated d-shell factors, Hobson's
theorem, the electrostatics of Gaussian charges, the 2-centre limit of the 3-centre form and
the normalisation PySCF applies to contracted shells."""
import math

import numpy as np
import pytest
import torch

from oracle import cint as oc


def _poly_eval(coef, l, v):
    return sum(c * v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** d for c, (a, b, d) in zip(coef, oc.cart_list(l)))


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4, 5])
def test_solid_harmonics_orthonormal_and_harmonic(l):
    from scipy.integrate import lebedev_rule
    x, w = lebedev_rule(29)  # exact for polynomials of degree <= 29 on the sphere
    v = x.T
    T = oc.c2s(l)
    Y = np.stack([_poly_eval(T[m], l, v) for m in range(2 * l + 1)])
    G = (Y * w) @ Y.T
    np.testing.assert_allclose(G, np.eye(2 * l + 1), atol=1e-12)
    # harmonic: the Laplacian of every row polynomial vanishes
    for m in range(2 * l + 1):
        lap = {}
        for c, (a, b, d) in zip(T[m], oc.cart_list(l)):
            for k, (aa, bb, dd) in enumerate([(a - 2, b, d), (a, b - 2, d), (a, b, d - 2)]):
                p = (a, b, d)[k]
                if p >= 2:
                    lap[(aa, bb, dd)] = lap.get((aa, bb, dd), 0.0) + c * p * (p - 1)
        assert max((abs(v) for v in lap.values()), default=0.0) < 1e-12


def test_libcint_d_factors():
    # libcint's d functions: xy, yz, z^2, xz, x^2 - y^2 with these factors
    T = oc.c2s(2)
    idx = {c: i for i, c in enumerate(oc.cart_list(2))}
    assert T[0, idx[(1, 1, 0)]] == pytest.approx(1.092548430592079070, abs=1e-12)
    assert T[1, idx[(0, 1, 1)]] == pytest.approx(1.092548430592079070, abs=1e-12)
    assert T[2, idx[(0, 0, 2)]] == pytest.approx(0.630783130505040012, abs=1e-12)
    assert T[2, idx[(2, 0, 0)]] == pytest.approx(-0.315391565252520002, abs=1e-12)
    assert T[3, idx[(1, 0, 1)]] == pytest.approx(1.092548430592079070, abs=1e-12)
    assert T[4, idx[(2, 0, 0)]] == pytest.approx(0.546274215296039535, abs=1e-12)
    assert T[4, idx[(0, 2, 0)]] == pytest.approx(-0.546274215296039535, abs=1e-12)
    np.testing.assert_allclose(oc.c2s(1), 0.488602511902919921 * np.eye(3), atol=1e-12)


@pytest.mark.parametrize("l", [1, 2, 3])
def test_hobson_theorem(l):
    """Y(d/dA) e^{-a |r-A|^2} = (2a)^l Y(r - A) e^{-a |r-A|^2} for each solid harmonic Y."""
    rng = np.random.default_rng(l)
    a = 0.7
    r = rng.standard_normal(3)
    T = oc.c2s(l)
    A = [torch.tensor(float(v), dtype=torch.float64, requires_grad=True) for v in rng.standard_normal(3)]
    f = torch.exp(-a * sum((float(r[i]) - A[i]) ** 2 for i in range(3)))
    for m in range(2 * l + 1):
        lhs = 0.0
        for c, (px, py, pz) in zip(T[m], oc.cart_list(l)):
            g = f
            for var, k in ((A[0], px), (A[1], py), (A[2], pz)):
                for _ in range(k):
                    g = torch.autograd.grad(g, var, create_graph=True)[0]
            lhs += c * float(g)
        d = r - np.array([float(x) for x in A])
        rhs = (2 * a) ** l * _poly_eval(T[m], l, d[None, :])[0] * float(f)
        assert lhs == pytest.approx(rhs, rel=1e-10, abs=1e-14)


@pytest.mark.parametrize("R", [1e-3, 0.4, 1.7, 6.0, 40.0])
def test_gaussian_charges_electrostatics(R):
    """Two normalised Gaussian charges (a/pi)^(3/2) e^{-a r^2}: interaction erf(sqrt(rho) R) / R
    (Gauss's law), rho = a g / (a + g)."""
    a, g = 0.9, 2.3
    A = torch.zeros(3, dtype=torch.float64)
    C = torch.tensor([0.0, R * 0.6, R * 0.8], dtype=torch.float64)
    v = float(oc.s2c(A, C, a, g)) * (a / math.pi) ** 1.5 * (g / math.pi) ** 1.5
    rho = a * g / (a + g)
    assert v == pytest.approx(math.erf(math.sqrt(rho) * R) / R, rel=1e-12)


def test_three_centre_tends_to_two_centre():
    A = torch.tensor([0.1, -0.3, 0.2], dtype=torch.float64)
    B = torch.tensor([1.0, 0.4, -0.7], dtype=torch.float64)
    C = torch.tensor([-0.5, 0.9, 0.3], dtype=torch.float64)
    two = float(oc.s2c(A, C, 1.1, 0.6))
    three = float(oc.s3c(A, B, C, 1.1, 1e-9, 0.6))
    assert three == pytest.approx(two, rel=1e-7)


def test_contracted_shell_normalised():
    """A d shell through make_env (PySCF's normalisation) integrates to 1 on a radial x Lebedev
    grid; so does a generally contracted s shell."""
    from scipy.integrate import lebedev_rule
    from hesship.cint import make_env
    basis = {"X": [(2, [[1.9, 0.3], [0.45, 0.8]]), (0, [[5.0, 0.4, 0.0], [1.2, 0.6, 0.2], [0.3, 0.1, 0.9]])]}
    atm, bas, env = make_env([("X", (0.0, 0.0, 0.0))], basis)
    xs, ws = lebedev_rule(29)
    # radial: Gauss-Chebyshev (second kind) mapped to r in (0, inf)
    n = 200
    t = np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    wt = np.pi / (n + 1) * np.sin(np.arange(1, n + 1) * np.pi / (n + 1)) ** 2 / np.sqrt(1 - t ** 2)
    r = (1 + t) / (1 - t)
    wr = wt * 2 / (1 - t) ** 2 * r ** 2
    pts = (r[:, None, None] * xs.T[None, :, :]).reshape(-1, 3)
    wts = (wr[:, None] * ws[None, :]).reshape(-1)
    shells = []
    for b in bas:
        l, nprim, nctr, pe, pc = b[1], b[2], b[3], b[5], b[6]
        ex = env[pe:pe + nprim]
        for c in range(nctr):
            shells.append((np.zeros(3), l, ex, env[pc + nprim * c:pc + nprim * (c + 1)]))
    ao = oc.eval_gto(shells, pts)[:, :, 0]
    norms = (ao ** 2 * wts[:, None]).sum(0)
    np.testing.assert_allclose(norms, 1.0, rtol=1e-8)


def test_golden_blocks_are_the_oracles():
    """The committed golden blocks (tests/golden/make_cint_golden.py) equal a fresh oracle
    evaluation (first two blocks of every kind)."""
    import os
    from hesship.cint import make_env
    from tests import cint_basis as cb
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cint_blocks.npz"))
    sa = cb.shells_of(*make_env(cb.ATOMS, cb.AO))
    sc = cb.shells_of(*make_env(cb.ATOMS, cb.AUX))
    for kind, maxl in cb.KINDS3.items():
        for n, (a, b, c) in enumerate(cb.blocks3(sa, sc, maxl)[:2]):
            ref = oc.block3c([a[1], b[1], c[1]], [(s[2], s[3], s[4]) for s in (a, b, c)], kind)
            np.testing.assert_allclose(g[f"3c_{kind}_{n}"], ref, rtol=1e-13, atol=1e-15)
    for kind, maxl in cb.KINDS2.items():
        for n, (a, c) in enumerate(cb.blocks2(sc, maxl)[:2]):
            ref = oc.block2c([a[1], c[1]], [(s[2], s[3], s[4]) for s in (a, c)], kind)
            np.testing.assert_allclose(g[f"2c_{kind}_{n}"], ref, rtol=1e-13, atol=1e-15)


# ---- one-electron integrals -----------------------------------------------------------------

def _norm_shell(l, a):
    from hesship.cint import make_env
    atm, bas, env = make_env([("X", (0.0, 0.0, 0.0))], {"X": [(l, [[a, 1.0]])]})
    return (l, [a], [float(env[bas[0][6]])])


@pytest.mark.parametrize("l", [0, 1, 2])
def test_one_electron_closed_forms(l):
    """Normalised e^{-a r^2} r^l Y_lm: <a|a> = 1, <a| -1/2 nabla^2 |a> = (2l + 3) a / 2, and for s
    <a| 1/r |a> = 2 sqrt(2a / pi) (origin at the centre)."""
    a = 0.83
    sh = _norm_shell(l, a)
    cen = [np.zeros(3), np.zeros(3)]
    np.testing.assert_allclose(oc.block1e("ovlp", cen, [sh, sh])[..., 0], np.eye(2 * l + 1), atol=1e-12)
    np.testing.assert_allclose(oc.block1e("kin", cen, [sh, sh])[..., 0], (2 * l + 3) * a / 2 * np.eye(2 * l + 1),
                               atol=1e-12)
    if l == 0:
        assert oc.block1e("rinv", cen, [sh, sh], C=np.zeros(3))[0, 0, 0] == pytest.approx(
            2 * math.sqrt(2 * a / math.pi), rel=1e-12)


def _grid(n_rad=160, order=41):
    from scipy.integrate import lebedev_rule
    xs, ws = lebedev_rule(order)
    t = np.cos(np.arange(1, n_rad + 1) * np.pi / (n_rad + 1))
    wt = np.pi / (n_rad + 1) * np.sin(np.arange(1, n_rad + 1) * np.pi / (n_rad + 1)) ** 2 / np.sqrt(1 - t ** 2)
    r = (1 + t) / (1 - t) * 1.2
    wr = wt * 2 * 1.2 / (1 - t) ** 2 * r ** 2
    pts = (r[:, None, None] * xs.T[None, :, :]).reshape(-1, 3)
    return pts, (wr[:, None] * ws[None, :]).reshape(-1)


@pytest.mark.parametrize("kind", ["plain", "ip1", "ipip1", "ipvip1"])
def test_one_electron_blocks_against_quadrature(kind):
    """A d shell at the origin against a contracted p shell 1.3 bohr away: overlap, kinetic (as
    1/2 int grad f . grad g) and 1/r at the origin, each kind from AO values and their
    derivatives on a radial x Lebedev grid centred at the origin (an independent route)."""
    A = np.zeros(3)
    B = np.array([0.5, -0.9, 0.8])
    sa = (2, [1.1], [_norm_shell(2, 1.1)[2][0]])
    sb = (1, [0.9, 0.35], [0.6, 0.5])
    pts, w = _grid()
    d = 3 if kind == "ipip1" else 2
    fa = oc.eval_gto([(A, *sa)], pts, deriv=d)   # [ng, 5, ncomp]
    fb = oc.eval_gto([(B, *sb)], pts, deriv=2)   # [ng, 3, ncomp]
    comp = {c: i for i, c in enumerate(oc.cart_list(0) + oc.cart_list(1) + oc.cart_list(2) + oc.cart_list(3))}

    def dd(f, dirs):
        e = [0, 0, 0]
        for x in dirs:
            e[x] += 1
        return f[:, :, comp[tuple(e)]]
    rinv = 1.0 / np.linalg.norm(pts, axis=1)
    kf = oc.KINDS1[kind]
    nq = 3 ** len(kf)
    for op in ("ovlp", "kin", "rinv"):
        ref = np.zeros((5, 3, nq))
        for q in range(nq):
            dirs = [(q // 3, q % 3)[i] if len(kf) == 2 else q for i in range(len(kf))]
            da = [x for f, x in zip(kf, dirs) if f == 0]
            db = [x for f, x in zip(kf, dirs) if f == 1]
            if op == "kin":
                val = sum(np.einsum("g,ga,gb->ab", 0.5 * w, dd(fa, da + [e]), dd(fb, db + [e])) for e in range(3))
            else:
                wt = w * (rinv if op == "rinv" else 1.0)
                val = np.einsum("g,ga,gb->ab", wt, dd(fa, da), dd(fb, db))
            ref[:, :, q] = val
        got = oc.block1e(op, [A, B], [sa, sb], kind, C=np.zeros(3))
        np.testing.assert_allclose(got, ref, atol=2e-7 * np.abs(ref).max(), err_msg=f"{op} {kind}")
