"""Harmonic analysis of a Cartesian Hessian (host FP64, library entry points hh_vib_*).

Mirrors src/hessian/vib.rs: `harmonic_analysis(hess, geom, mass, project_trans, project_rot)`
returns the VibInfo fields (vib.rs:221-256); `rotation_const`, `get_rotor_type` and
`get_tr_space` are the reference's helpers of the same names.  Arrays are NumPy, float64,
column-major as in the reference (geom [3, natm], hess [3 natm, 3 natm]).
"""
import ctypes as C
from types import SimpleNamespace

import numpy as np

from ._lib import check, lib

ROTOR = {0: "ATOM", 1: "LINEAR", 2: "REGULAR"}
TRV = {0: "V", 1: "TR", 2: "-"}


class _VibOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("omega", "imag", "q", "w", "x", "degeneracy", "trv", "mu", "k", "dq0",
                                           "qtp0", "xtp0", "theta_vib")]


def _f64(x):
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def harmonic_analysis(hess, geom, mass, project_trans=True, project_rot=True):
    geom, mass, hess = _f64(geom), _f64(mass).ravel(), _f64(hess)
    natm = mass.shape[0]
    n = 3 * natm
    if geom.shape != (3, natm) or hess.shape != (n, n):
        raise ValueError("geom must be [3, natm] and hess [3 natm, 3 natm]")
    r = SimpleNamespace(ndof=n)
    for name in ("omega", "mu", "k", "dq0", "qtp0", "xtp0", "theta_vib"):
        setattr(r, name, np.zeros(n))
    for name in ("q", "w", "x"):
        setattr(r, name, np.zeros((n, n), order="F"))
    r.imag = np.zeros(n, dtype=np.int32)
    r.trv_code = np.zeros(n, dtype=np.int32)
    r.degeneracy = np.zeros(n, dtype=np.int64)
    out = _VibOut(*[_ptr(a) for a in (r.omega, r.imag, r.q, r.w, r.x, r.degeneracy, r.trv_code, r.mu, r.k, r.dq0,
                                      r.qtp0, r.xtp0, r.theta_vib)])
    check(lib().hh_vib_harmonic_analysis(_ptr(hess), _ptr(geom), _ptr(mass), natm, int(project_trans),
                                         int(project_rot), C.byref(out)))
    r.imag = r.imag.astype(bool)
    r.trv = [TRV[int(t)] for t in r.trv_code]
    r.vib_indices = [i for i, t in enumerate(r.trv) if t == "V"]
    return r


def rotation_const(mass, geom_c, unit="GHz"):
    mass, geom_c = _f64(mass).ravel(), _f64(geom_c)
    out = np.zeros(3)
    u = {"ghz": 0, "wavenumber": 1}[unit.lower()]
    check(lib().hh_vib_rotation_const(_ptr(mass), _ptr(geom_c), mass.shape[0], u, _ptr(out)))
    return out


def get_rotor_type(rc_ghz):
    rc = _f64(rc_ghz).ravel()
    t = C.c_int32()
    check(lib().hh_vib_rotor_type(_ptr(rc), C.byref(t)))
    return ROTOR[t.value]


def get_tr_space(mass, geom, space="TR"):
    mass, geom = _f64(mass).ravel(), _f64(geom)
    n = 3 * mass.shape[0]
    tr = np.zeros((n, 6), order="F")
    nc = C.c_int64()
    bits = (1 if "T" in space else 0) | (2 if "R" in space else 0)
    check(lib().hh_vib_tr_space(_ptr(mass), _ptr(geom), mass.shape[0], bits, _ptr(tr), C.byref(nc)))
    return tr[:, :nc.value].copy()
