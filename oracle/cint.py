"""CPU oracle (FP64) of the Gaussian-basis integral and AO-evaluation engine: test
infrastructure only (imported by tests/, never by the library or the timed bench paths).

The reference takes these values from libcint (third-party, absent here: `libcint = "0.3"`,
Cargo.toml:7; called at cint_handling.rs:20-92, nimatmul.rs:94-99), so there are no reference
vectors; the oracle is built from the mathematics, by a route independent of the device code:

- real solid harmonics: fitted by least squares to scipy's complex spherical harmonics
  (r^l Y_lm at random points; real forms sqrt(2) (-1)^m Re / Im Y_l^{|m|}, libcint's order
  m = -l..l, p as x, y, z), not from the closed form the device uses;
- AO values and their derivatives: the function written directly, derivatives by torch autograd;
- Coulomb integrals: only the s-type closed forms (ss|s) = 2 pi^(5/2) / (p g sqrt(p + g))
  exp(-ab/p |A - B|^2) F_0(rho |P - C|^2) and (s|s) = 2 pi^(5/2) / (a g sqrt(a + g)) F_0(..),
  with F_0(T) = sqrt(pi / 4T) erf(sqrt T) (its Taylor series near 0); higher angular momenta by
  Hobson's theorem (Y(d/dA) e^{-a |r-A|^2} = (2a)^l Y(r - A) e^{..} for a harmonic polynomial Y)
  as derivatives of the s-type integral with respect to the centres, taken with torch autograd;
  libcint's ip on a function is minus the derivative with respect to that function's centre.
  Cost grows with the total derivative order: use at most ~5.
"""
import math
from functools import lru_cache

import numpy as np
import torch


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cart_list(l):
    """libcint's Cartesian order: x^l, x^{l-1} y, x^{l-1} z, x^{l-2} y^2, ..."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


@lru_cache(maxsize=None)
def c2s(l):
    """[(2l+1), ncart(l)]: rows = real solid harmonics r^l Y_lm (unit-sphere normalised) as
    combinations of the Cartesian monomials, libcint's row order."""
    from scipy.special import sph_harm_y
    rng = np.random.default_rng(100 + l)
    npt = 8 * ncart(l) + 20
    v = rng.standard_normal((npt, 3))
    r = np.linalg.norm(v, axis=1) * rng.uniform(0.5, 1.5, npt)  # random radii: monomials independent
    v = v / np.linalg.norm(v, axis=1)[:, None] * r[:, None]
    theta = np.arccos(np.clip(v[:, 2] / r, -1, 1))
    phi = np.arctan2(v[:, 1], v[:, 0])
    mono = np.stack([v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** c for a, b, c in cart_list(l)], axis=1)
    order = list(range(-l, l + 1)) if l != 1 else [1, -1, 0]
    rows = []
    for m in order:
        y = sph_harm_y(l, abs(m), theta, phi)
        if m > 0:
            yr = math.sqrt(2.0) * (-1) ** m * y.real
        elif m < 0:
            yr = math.sqrt(2.0) * (-1) ** abs(m) * y.imag
        else:
            yr = y.real
        coef, *_ = np.linalg.lstsq(mono, r ** l * yr, rcond=None)
        rows.append(coef)
    return np.array(rows)


def common_fac(l):
    return {0: 0.282094791773878143, 1: 0.488602511902919921}.get(l, 1.0)


def gto_norm(l, a):
    """1 / sqrt(int_0^inf r^(2l+2) e^(-2 a r^2) dr)"""
    n = 2 * l + 2
    return 1.0 / math.sqrt(math.gamma((n + 1) / 2) / (2.0 * (2 * a) ** ((n + 1) / 2)))


# ---- AO values ----------------------------------------------------------------------------

def eval_gto(shells, coords, deriv=0, cart=False):
    """shells: list of (centre [3], l, exps [nprim], coefs [nprim]) (one contraction each, in AO
    order); coords [ng, 3].  Returns [ng, nao, ncomp] like the device (libcint's component order)."""
    x = torch.tensor(coords[:, 0], dtype=torch.float64, requires_grad=True)
    y = torch.tensor(coords[:, 1], dtype=torch.float64, requires_grad=True)
    z = torch.tensor(coords[:, 2], dtype=torch.float64, requires_grad=True)
    comps = [(dx, dy, dz) for n in range(deriv + 1) for (dx, dy, dz) in cart_list(n)]
    cols = []
    for (A, l, ex, co) in shells:
        rx, ry, rz = x - A[0], y - A[1], z - A[2]
        r2 = rx * rx + ry * ry + rz * rz
        rad = sum(float(c) * torch.exp(-float(a) * r2) for a, c in zip(ex, co))
        mono = [rx ** a * ry ** b * rz ** c for a, b, c in cart_list(l)]
        if cart:
            fns = [common_fac(l) * m * rad for m in mono]
        else:
            T = c2s(l)
            fns = [sum(float(T[m, i]) * mono[i] for i in range(ncart(l))) * rad for m in range(2 * l + 1)]
        for f in fns:
            out = []
            for dx, dy, dz in comps:
                g = f
                for var, k in ((x, dx), (y, dy), (z, dz)):
                    for _ in range(k):
                        g = torch.autograd.grad(g.sum(), var, create_graph=True)[0]
                out.append(g.detach().numpy())
            cols.append(np.stack(out, axis=1))
    return np.stack(cols, axis=1)


# ---- Coulomb integrals --------------------------------------------------------------------

def _f0(T):
    if float(T.detach()) < 0.5:
        return sum((-T) ** k / float(math.factorial(k) * (2 * k + 1)) for k in range(30))
    s = torch.sqrt(T)
    return 0.5 * math.sqrt(math.pi) * torch.erf(s) / s


def s3c(A, B, C, a, b, g):
    """(s_a s_b | s_c) of unnormalised primitives, torch function of the centres"""
    p = a + b
    P = (a * A + b * B) / p
    rho = p * g / (p + g)
    AB = A - B
    PC = P - C
    pref = 2.0 * math.pi ** 2.5 / (p * g * math.sqrt(p + g))
    return pref * torch.exp(-a * b / p * (AB * AB).sum()) * _f0(rho * (PC * PC).sum())


def s2c(A, C, a, g):
    rho = a * g / (a + g)
    AC = A - C
    return 2.0 * math.pi ** 2.5 / (a * g * math.sqrt(a + g)) * _f0(rho * (AC * AC).sum())


class _Deriv:
    """Mixed partial derivatives of a scalar torch function of 9 centre coordinates, memoised
    by the (sorted) multiset of variables."""

    def __init__(self, fn, vars_):
        self.vars = vars_
        self.memo = {(): fn()}

    def get(self, key):
        key = tuple(sorted(key))
        if key in self.memo:
            return self.memo[key]
        prev = self.get(key[:-1])
        if prev.requires_grad:
            d = torch.autograd.grad(prev, self.vars[key[-1]], create_graph=True, allow_unused=True)[0]
            if d is None:
                d = torch.zeros((), dtype=torch.float64)
        else:
            d = torch.zeros((), dtype=torch.float64)
        self.memo[key] = d
        return d


# derivative directions (t fast for one derivative; 9 components c = 3 t + s, t on the earlier
# function) per integral kind: list of (function index 0 = a, 1 = b, 2 = c) per ip
KINDS3 = {"plain": [], "ip1": [0], "ip2": [2], "ipip1": [0, 0], "ipvip1": [0, 1], "ip1ip2": [0, 2], "ipip2": [2, 2]}
KINDS2 = {"plain": [], "ip1": [0], "ipip1": [0, 0], "ip1ip2": [0, 2]}


def _shell_block(centres, shells, kind_fns, two_c):
    """Integrals over one (a, b, c) shell triple (b absent for 2-centre): array
    [2la+1, (2lb+1), 2lc+1, ncomp]."""
    ncomp = 3 ** len(kind_fns)
    la, lc = shells[0][1], shells[-1][1]
    lb = 0 if two_c else shells[1][1]
    Ta, Tc = c2s(la), c2s(lc)
    Tb = np.ones((1, 1)) if two_c else c2s(lb)
    ca, cc = cart_list(la), cart_list(lc)
    cb = [(0, 0, 0)] if two_c else cart_list(lb)
    out = np.zeros((2 * la + 1, Tb.shape[0], 2 * lc + 1, ncomp))
    prims = [list(zip(s[2], s[3])) for s in shells]
    if two_c:
        prims = [prims[0], [(0.0, 1.0)], prims[1]]
    # every primitive combination in one scalar: sum_p w_p (s s|s)_p, w_p = contraction weights
    # times Hobson's (2 alpha)^-l (the derivative operators are linear)
    combos = [(ea, wa, eb, wb, ec, wc) for (ea, wa) in prims[0] for (eb, wb) in prims[1] for (ec, wc) in prims[2]]
    vs = [torch.tensor(float(v), dtype=torch.float64, requires_grad=True) for f in range(3) for v in centres[f]]

    def fn():
        Av, Bv, Cv = torch.stack(vs[0:3]), torch.stack(vs[3:6]), torch.stack(vs[6:9])
        tot = 0.0
        for ea, wa, eb, wb, ec, wc in combos:
            w = wa * wb * wc * (2 * ea) ** (-la) * (1.0 if two_c else (2 * eb) ** (-lb)) * (2 * ec) ** (-lc)
            tot = tot + w * (s2c(Av, Cv, ea, ec) if two_c else s3c(Av, Bv, Cv, ea, eb, ec))
        return tot
    D = _Deriv(fn, vs)
    sign = (-1) ** len(kind_fns)
    for q in range(ncomp):
        dirs = [(q // 3, q % 3)[i] if len(kind_fns) == 2 else q for i in range(len(kind_fns))]
        extra = [3 * f + d for f, d in zip(kind_fns, dirs)]
        for ia, (ax, ay, az) in enumerate(ca):
            for ib, (bx, by, bz) in enumerate(cb):
                for ic, (cx, cy, cz) in enumerate(cc):
                    key = [0] * ax + [1] * ay + [2] * az + [3] * bx + [4] * by + [5] * bz + \
                          [6] * cx + [7] * cy + [8] * cz + extra
                    v = float(D.get(key).detach()) * sign
                    if v == 0.0:
                        continue
                    out[:, :, :, q] += v * np.einsum("a,b,c->abc", Ta[:, ia], Tb[:, ib], Tc[:, ic])
    return out


def int3c2e(mol, aux, kind="plain"):
    """mol / aux: (centres [natm, 3], shells = list of (atom, l, exps, coefs)); spherical.
    Returns [nao, nao, naux, comp] (comp 9 as [3, 3] flattened with the later function's
    direction fastest... as the device: c = 3 t + s)."""
    kf = KINDS3[kind]
    ao = np.cumsum([0] + [2 * s[1] + 1 for s in mol[1]])
    ax = np.cumsum([0] + [2 * s[1] + 1 for s in aux[1]])
    out = np.zeros((ao[-1], ao[-1], ax[-1], 3 ** len(kf)))
    for i, si in enumerate(mol[1]):
        for j, sj in enumerate(mol[1]):
            for k, sk in enumerate(aux[1]):
                cen = [mol[0][si[0]], mol[0][sj[0]], aux[0][sk[0]]]
                blk = _shell_block(cen, [si, sj, sk], kf, False)
                out[ao[i]:ao[i + 1], ao[j]:ao[j + 1], ax[k]:ax[k + 1]] = blk
    return out


def int2c2e(aux, kind="plain"):
    kf = KINDS2[kind]
    kf = [0 if f == 0 else 2 for f in kf]
    ax = np.cumsum([0] + [2 * s[1] + 1 for s in aux[1]])
    out = np.zeros((ax[-1], ax[-1], 3 ** len(kf)))
    for i, si in enumerate(aux[1]):
        for k, sk in enumerate(aux[1]):
            cen = [aux[0][si[0]], aux[0][si[0]], aux[0][sk[0]]]
            blk = _shell_block(cen, [si, sk], kf, True)
            out[ax[i]:ax[i + 1], ax[k]:ax[k + 1]] = blk[:, 0]
    return out


def block3c(centres, shells, kind="plain"):
    """One (a, b, c) shell triple of int3c2e_<kind>: centres [3][3], shells [(l, exps, coefs)] x 3;
    returns [2la+1, 2lb+1, 2lc+1, comp]."""
    sh = [(None, l, ex, co) for (l, ex, co) in shells]
    return _shell_block(centres, sh, KINDS3[kind], False)


def block2c(centres, shells, kind="plain"):
    """One (a, c) shell pair of int2c2e_<kind>: centres [2][3], shells [(l, exps, coefs)] x 2;
    returns [2la+1, 2lc+1, comp]."""
    sh = [(None, l, ex, co) for (l, ex, co) in shells]
    return _shell_block([centres[0], centres[0], centres[1]], sh, KINDS2[kind], True)[:, 0]


# ---- one-electron integrals ---------------------------------------------------------------
# s-type closed forms of unnormalised primitives e^{-a |r-A|^2}, e^{-b |r-B|^2} (p = a + b,
# mu = ab / p, P = (aA + bB) / p):
#   overlap  (pi / p)^(3/2) e^{-mu |AB|^2}
#   kinetic  mu (3 - 2 mu |AB|^2) (pi / p)^(3/2) e^{-mu |AB|^2}       (<a| -1/2 nabla^2 |b>)
#   1 / |r - C|   2 pi / p e^{-mu |AB|^2} F_0(p |P - C|^2)
# higher angular momenta by Hobson's theorem as above (the operators do not depend on A, B).

def s1e(op, A, B, C, a, b):
    p = a + b
    mu = a * b / p
    AB = A - B
    ab2 = (AB * AB).sum()
    if op == "ovlp":
        return (math.pi / p) ** 1.5 * torch.exp(-mu * ab2)
    if op == "kin":
        return mu * (3.0 - 2.0 * mu * ab2) * (math.pi / p) ** 1.5 * torch.exp(-mu * ab2)
    if op == "rinv":
        P = (a * A + b * B) / p
        PC = P - C
        return 2.0 * math.pi / p * torch.exp(-mu * ab2) * _f0(p * (PC * PC).sum())
    raise ValueError(op)


KINDS1 = {"plain": [], "ip1": [0], "ipip1": [0, 0], "ipvip1": [0, 1]}


def block1e(op, centres, shells, kind="plain", C=None):
    """One (a, b) shell pair of int1e_<op> (op "ovlp", "kin" or "rinv" with origin C):
    centres [2][3], shells [(l, exps, coefs)] x 2 (coefficients as the engine takes them);
    returns [2la+1, 2lb+1, comp] (comp 9 = 3 t + s, t on a)."""
    kind_fns = KINDS1[kind]
    ncomp = 3 ** len(kind_fns)
    (la, exa, coa), (lb, exb, cob) = shells
    Ta, Tb = c2s(la), c2s(lb)
    out = np.zeros((2 * la + 1, 2 * lb + 1, ncomp))
    Cv = torch.tensor(np.zeros(3) if C is None else np.asarray(C, dtype=np.float64), dtype=torch.float64)
    vs = [torch.tensor(float(v), dtype=torch.float64, requires_grad=True) for f in range(2) for v in centres[f]]
    combos = [(ea, wa, eb, wb) for ea, wa in zip(exa, coa) for eb, wb in zip(exb, cob)]

    def fn():
        Av, Bv = torch.stack(vs[0:3]), torch.stack(vs[3:6])
        tot = 0.0
        for ea, wa, eb, wb in combos:
            tot = tot + wa * wb * (2 * ea) ** (-la) * (2 * eb) ** (-lb) * s1e(op, Av, Bv, Cv, ea, eb)
        return tot
    D = _Deriv(fn, vs)
    sign = (-1) ** len(kind_fns)
    for q in range(ncomp):
        dirs = [(q // 3, q % 3)[i] if len(kind_fns) == 2 else q for i in range(len(kind_fns))]
        extra = [3 * f + d for f, d in zip(kind_fns, dirs)]
        for ia, (ax, ay, az) in enumerate(cart_list(la)):
            for ib, (bx, by, bz) in enumerate(cart_list(lb)):
                key = [0] * ax + [1] * ay + [2] * az + [3] * bx + [4] * by + [5] * bz + extra
                v = float(D.get(key).detach()) * sign
                if v != 0.0:
                    out[:, :, q] += v * np.outer(Ta[:, ia], Tb[:, ib])
    return out
