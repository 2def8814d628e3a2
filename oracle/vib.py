"""Oracle: harmonic analysis (NumPy FP64).  Test infrastructure only.

Restates src/hessian/vib.rs (itself a port of psi4 qcdb/vib.py) with NumPy's SVD and
symmetric eigensolver, so the library's own Jacobi routines are checked against independent
numerics.  Pinned by the reference's hard-coded fixtures (tests/test_vib.rs: NH3 rotational
constants, rotor classification, TR-space shapes and orthonormality) and by Hessians built
from known normal modes (closed-form frequencies, reduced masses and force constants).
"""
import numpy as np

# CODATA 2014 constants of vib.rs:18-29
NA = 6.022140857e23
HARTREE2J = 4.35974465e-18
C = 299792458.0
BOHR2ANG = 0.52917721067
H = 6.62607004e-34
KB = 1.38064852e-23
AMU2KG = 1.66053904e-27


def uconv_cm1():
    """vib.rs:33-35."""
    return np.sqrt(NA * HARTREE2J * 1.0e19) / (2.0 * np.pi * C * BOHR2ANG)


def tr_raw(mass, geom, space="TR"):
    """The six idealised TR vectors [3 natm, n] (vib.rs:66-122)."""
    natm = len(mass)
    sm = np.repeat(np.sqrt(mass), 3)
    x, y, z = (np.repeat(geom[c], 3) for c in range(3))
    ux, uy, uz = (np.tile(np.eye(3)[c], natm) for c in range(3))
    cols = []
    if "T" in space:
        cols += [sm * ux, sm * uy, sm * uz]
    if "R" in space:
        cols += [sm * (y * uz - z * uy), sm * (z * ux - x * uz), sm * (x * uy - y * ux)]
    if not cols:
        cols = [np.zeros(3 * natm)]
    return np.stack(cols, axis=1)


def get_tr_space(mass, geom, space="TR", nrt=None):
    """vib.rs:59-140: left singular vectors; rank nrt, else above 3 natm * max(s) * eps."""
    u, s, _ = np.linalg.svd(tr_raw(mass, geom, space), full_matrices=False)
    if nrt is None:
        nrt = int((s > len(u) * s.max() * np.finfo(float).eps).sum())
    return u[:, :nrt]


def rotation_const(mass, geom_c, unit="GHz"):
    """vib.rs:152-197 (constants ascending in moment of inertia)."""
    rr = (geom_c * mass) @ geom_c.T
    im = np.eye(3) * np.trace(rr) - rr
    e = np.linalg.eigvalsh(im)
    e = np.where(np.abs(e) < 1e-9, 0.0, e)
    unit_im = AMU2KG * 5.2917721092e-11 ** 2
    unit_hz = 1.0545718001391127e-34 / (4.0 * np.pi * unit_im)
    conv = {"ghz": 1e-9, "wavenumber": 1.0 / C * 1e-2}[unit.lower()]
    with np.errstate(divide="ignore"):
        return np.where(np.abs(e) < 1e-30, np.inf, unit_hz / np.where(e == 0, 1.0, e) * conv)


def get_rotor_type(rc):
    """vib.rs:203-211."""
    if all(x > 1e8 for x in rc):
        return "ATOM"
    if rc[0] > 1e8 and abs(rc[1] - rc[2]) < 1e-3:
        return "LINEAR"
    return "REGULAR"


def harmonic_analysis(hess, geom, mass, project_trans=True, project_rot=True):
    """vib.rs:336-552; returns a dict of the VibInfo fields."""
    natm = len(mass)
    n = 3 * natm
    gc = geom - (geom * mass).sum(1, keepdims=True) / mass.sum()
    rotor = get_rotor_type(rotation_const(mass, gc, "GHz"))
    nrt = {"ATOM": 0, "LINEAR": 5, "REGULAR": 6}[rotor]
    if not project_trans:
        nrt = max(nrt, 3) - 3
    if not project_rot:
        nrot = {"ATOM": 0, "LINEAR": 2, "REGULAR": 3}[rotor]
        nrt = max(nrt, nrot) - nrot
    space = ("T" if project_trans else "") + ("R" if project_rot else "")
    tr = get_tr_space(mass, geom, space, nrt)
    P = np.eye(n) - tr @ tr.T
    smi = 1.0 / np.repeat(np.sqrt(mass), 3)
    mw = hess * smi[:, None] * smi[None, :]
    fc, q = np.linalg.eigh(P.T @ mw @ P)
    order = np.argsort(fc, kind="stable")
    fc, q = fc[order], q[:, order]
    for v in range(n):  # phase convention (vib.rs:283-306)
        a = np.abs(q[:, v])
        iext = int(np.nonzero(a.max() - a < 1e-2)[0][0])
        if q[iext, v] < 0:
            q[:, v] = -q[:, v]
    imag = fc < 0
    omega = np.sqrt(np.abs(fc)) * uconv_cm1()
    wre = np.where(imag, 0.0, omega)
    keys = np.rint(wre * 10.0).astype(np.int64)
    degeneracy = np.array([(keys == k).sum() for k in keys], dtype=np.int64)
    is_tr = np.zeros(n, dtype=bool)
    is_tr[np.argsort(np.abs(fc), kind="stable")[:nrt]] = True
    trv = ["TR" if is_tr[i] else ("-" if abs(fc[i]) < 1e-3 else "V") for i in range(n)]
    uconv_mdyne_a = 0.1 * (2.0 * np.pi * C) ** 2 / NA
    uconv_s = np.sqrt((C * (2.0 * np.pi * BOHR2ANG) ** 2) / (H * NA * 1.0e21))
    w = q * smi[:, None]
    mu = 1.0 / (w ** 2).sum(0)
    x = w * np.sqrt(mu)[None, :]
    k = mu * wre ** 2 * uconv_mdyne_a
    with np.errstate(divide="ignore"):
        dq = np.sqrt(wre) * uconv_s
        dx = np.sqrt(wre * mu) * uconv_s
        qtp0 = np.where(dq == 0, 0.0, 1.0 / np.where(dq == 0, 1.0, dq))
        xtp0 = np.where(dx == 0, 0.0, 1.0 / np.where(dx == 0, 1.0, dx))
    theta = wre * 100.0 * H * C / KB
    return dict(ndof=n, fc=fc, omega=omega, imag=imag, q=q, w=w, x=x, degeneracy=degeneracy, trv=trv, mu=mu, k=k,
                qtp0=qtp0, xtp0=xtp0, dq0=qtp0 / np.sqrt(2.0), theta_vib=theta, nrt=nrt, rotor=rotor)


def hessian_from_modes(mass, geom, lam, seed=0, linear=False):
    """Cartesian Hessian whose projected mass-weighted form has eigenvalues `lam` (len = 3 natm -
    nrt) on random orthonormal directions orthogonal to the TR space, plus the matching modes:
    returns (hess, modes [3 natm, nvib] mass-weighted).  The closed form the checks use."""
    n = 3 * len(mass)
    tr = get_tr_space(mass, geom, "TR", 5 if linear else 6)
    rng = np.random.default_rng(seed)
    Z = rng.standard_normal((n, len(lam)))
    Z -= tr @ (tr.T @ Z)
    Q, _ = np.linalg.qr(Z)
    Q -= tr @ (tr.T @ Q)
    Q, _ = np.linalg.qr(Q)
    hmw = Q @ np.diag(lam) @ Q.T
    sm = np.repeat(np.sqrt(mass), 3)
    return hmw * sm[:, None] * sm[None, :], Q
