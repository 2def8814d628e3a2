"""Gaussian-basis integrals and AO values on the device (the libcint calls of the reference's
Hessian path: CInt::eval_gto, nimatmul.rs:94-99; int2c2e / int3c2e and their derivatives,
cint_handling.rs:20-92).

CInt holds a basis in libcint's (atm, bas, env) format.  `CInt.from_basis` builds those arrays
from atoms and a basis given per element as [(l, [[exp, c_1, c_2, ...], ...]), ...] (general
contractions allowed), with the normalisation PySCF applies before calling libcint (each
primitive scaled by 1 / sqrt(int r^(2l+2) e^(-2 a r^2) dr), each contraction normalised).
"""
import ctypes as C
import math

import numpy as np
import torch

from ._lib import check, lib
from .tensors import empty_f, stream_of

KIND = {"plain": 0, "ip1": 1, "ip2": 2, "ipip1": 3, "ipvip1": 4, "ip1ip2": 5, "ipip2": 6}
NCOMP = {"plain": 1, "ip1": 3, "ip2": 3, "ipip1": 9, "ipvip1": 9, "ip1ip2": 9, "ipip2": 9}
# libcint one-electron integral names -> (operator, kind) of hh_cint_int1e
INT1E_OP = {"ovlp": 0, "kin": 1, "nuc": 2, "rinv": 3}
INT1E = {}
for _op in INT1E_OP:
    INT1E[f"int1e_{_op}"] = (_op, "plain")
    INT1E[f"int1e_ip{_op}"] = (_op, "ip1")
    INT1E[f"int1e_ipip{_op}"] = (_op, "ipip1")
    INT1E[f"int1e_ip{_op}ip"] = (_op, "ipvip1")
ELEMENTS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"]


def charge_of(sym):
    """Nuclear charge of an element symbol (a label such as "H1" or "O-a" reads as its element)."""
    base = "".join(ch for ch in str(sym) if ch.isalpha())
    for n in (2, 1):
        if base[:n].capitalize() in ELEMENTS:
            return ELEMENTS.index(base[:n].capitalize())
    return 0


def _gint(n, a):
    """int_0^inf r^n e^(-a r^2) dr"""
    return math.gamma((n + 1) / 2) / (2.0 * a ** ((n + 1) / 2))


def make_env(atoms, basis):
    """atoms: [(symbol, (x, y, z) in bohr)]; basis: {symbol: [(l, [[exp, c1, ...], ...]), ...]}.
    Returns (atm int32 [natm, 6], bas int32 [nbas, 8], env float64)."""
    env = [0.0] * 20  # libcint reserves the first 20 slots
    atm, bas = [], []
    for ia, (sym, xyz) in enumerate(atoms):
        atm.append([charge_of(sym), len(env), 1, 0, 0, 0])
        env.extend(float(v) for v in xyz)
    for ia, (sym, _) in enumerate(atoms):
        for l, rows in basis[sym]:
            rows = np.asarray(rows, dtype=np.float64)
            es, cs = rows[:, 0], rows[:, 1:].copy()
            nprim, nctr = cs.shape
            cs *= np.array([1.0 / math.sqrt(_gint(2 * l + 2, 2 * e)) for e in es])[:, None]
            ee = np.array([[_gint(2 * l + 2, a + b) for b in es] for a in es])
            cs /= np.sqrt(np.einsum("pi,pq,qi->i", cs, ee, cs))[None, :]
            pexp = len(env)
            env.extend(es.tolist())
            pcoef = len(env)
            env.extend(cs.reshape(-1, order="F").tolist())
            bas.append([ia, l, nprim, nctr, 0, pexp, pcoef, 0])
    return (np.ascontiguousarray(atm, dtype=np.int32), np.ascontiguousarray(bas, dtype=np.int32),
            np.ascontiguousarray(env, dtype=np.float64))


class CInt:
    """A basis on the device (libcint's CInt role).  atm / bas / env as libcint takes them."""

    def __init__(self, atm, bas, env, cart=False):
        self.atm = np.ascontiguousarray(atm, dtype=np.int32)
        self.bas = np.ascontiguousarray(bas, dtype=np.int32)
        self.env = np.ascontiguousarray(env, dtype=np.float64)
        self.cart = bool(cart)
        h = C.c_void_p()
        check(lib().hh_cint_create(self.atm.ctypes.data, len(self.atm), self.bas.ctypes.data, len(self.bas),
                                   self.env.ctypes.data, len(self.env), int(self.cart), C.byref(h)))
        self._h = h
        n = C.c_int64()
        loc = (C.c_int64 * (len(self.bas) + 1))()
        check(lib().hh_cint_nao(h, C.byref(n), loc))
        self.nao = n.value
        self.ao_loc = np.array(loc[:], dtype=np.int64)

    @classmethod
    def from_basis(cls, atoms, basis, cart=False):
        return cls(*make_env(atoms, basis), cart=cart)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().hh_cint_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def nbas(self):
        return len(self.bas)

    def atom_coords(self):
        return np.array([self.env[a[1]:a[1] + 3] for a in self.atm])

    def aoslice_by_atom(self):
        """[natm, 2] AO ranges per atom (shells are grouped by atom in from_basis order)."""
        natm = len(self.atm)
        out = np.zeros((natm, 2), dtype=np.int64)
        for A in range(natm):
            sh = np.nonzero(self.bas[:, 0] == A)[0]
            out[A] = (self.ao_loc[sh.min()], self.ao_loc[sh.max() + 1]) if len(sh) else (0, 0)
        return out

    def eval_gto(self, coords, deriv=0):
        """[ng, nao, ncomp] (column-major device tensor), coords [ng, 3] (device or host)."""
        c = torch.as_tensor(coords, dtype=torch.float64)
        dev = c.device if c.is_cuda else torch.device("cuda", torch.cuda.current_device())
        c = c.to(dev).contiguous()
        ng = c.shape[0]
        ncomp = (deriv + 1) * (deriv + 2) * (deriv + 3) // 6
        out = empty_f((ng, self.nao, ncomp), dev)
        check(lib().hh_cint_eval_gto(self._h, deriv, c.data_ptr(), ng, out.data_ptr(), stream_of(dev)))
        return out

    def int2c2e(self, kind="plain", device=None):
        """[naux, naux(, 3 | 3, 3)] (hess_intor reshape of the 9 components)."""
        dev = device or torch.device("cuda", torch.cuda.current_device())
        nc = NCOMP[kind]
        out = empty_f((self.nao, self.nao, nc), dev)
        check(lib().hh_cint_int2c2e(self._h, KIND[kind], out.data_ptr(), stream_of(dev)))
        return _shape(out, nc)

    def int3c2e(self, aux, kind="plain", shls=None, device=None):
        """[nao, nao, n01(, 3 | 3, 3)] over the auxiliary shells shls = (sh0, sh1) (all if None)."""
        dev = device or torch.device("cuda", torch.cuda.current_device())
        sh0, sh1 = shls if shls is not None else (0, aux.nbas)
        nc = NCOMP[kind]
        n01 = int(aux.ao_loc[sh1] - aux.ao_loc[sh0])
        out = empty_f((self.nao, self.nao, n01, nc), dev)
        check(lib().hh_cint_int3c2e(self._h, aux._h, KIND[kind], sh0, sh1, out.data_ptr(), stream_of(dev)))
        return _shape(out, nc)


    def atom_charges(self):
        return self.atm[:, 0].astype(np.float64)

    def intor(self, name, rinv_atom=None, device=None):
        """One-electron integrals by libcint name (int1e_ovlp, int1e_kin, int1e_nuc, int1e_rinv and
        their ip / ipip / ip..ip forms): [nao, nao(, 3 | 3, 3)]; int1e_*rinv* at nucleus rinv_atom
        (no charge: the caller scales by -Z as with_rinv_at_nucleus users do)."""
        if name not in INT1E:
            raise ValueError(f"unknown one-electron integral {name!r}")
        op, kind = INT1E[name]
        if op == "rinv" and rinv_atom is None:
            raise ValueError("rinv integrals need rinv_atom")
        dev = device or torch.device("cuda", torch.cuda.current_device())
        nc = NCOMP[kind]
        out = empty_f((self.nao, self.nao, nc), dev)
        check(lib().hh_cint_int1e(self._h, INT1E_OP[op], KIND[kind], -1 if rinv_atom is None else int(rinv_atom),
                                  out.data_ptr(), stream_of(dev)))
        return _shape(out, nc)

    def int3c2e_s2ij(self, aux, shls=None, device=None):
        """[nao_tp, n01]: the packed upper triangle (mu <= nu) of int3c2e ("s2ij", hess_r.rs:1731)."""
        dev = device or torch.device("cuda", torch.cuda.current_device())
        sh0, sh1 = shls if shls is not None else (0, aux.nbas)
        n01 = int(aux.ao_loc[sh1] - aux.ao_loc[sh0])
        out = empty_f((self.nao * (self.nao + 1) // 2, n01), dev)
        check(lib().hh_cint_int3c2e_s2ij(self._h, aux._h, sh0, sh1, out.data_ptr(), stream_of(dev)))
        return out


def _shape(t, nc):
    if nc == 1:
        return t[..., 0]
    if nc == 3:
        return t
    return _nine(t)


def _nine(t):
    """9 components (c = 3 t + s, t on the earlier function) as [..., 3 (s), 3 (t)] column-major,
    i.e. the same memory as hess_intor's reshape (cint_handling.rs:20-44)."""
    # t is column-major [.., 9]; element (.., s + 3 t) -> (.., s, t) keeps the memory order
    return torch.as_strided(t, t.shape[:-1] + (3, 3), t.stride()[:-1] + (t.stride()[-1], 3 * t.stride()[-1]))


class Int3cGenerator:
    """The skeleton's 3-centre integral generator (generator_hess_intor_j3c_by_aux,
    cint_handling.rs:80-92) on the device engine: pass `.fn` and `.ctx` to the skeleton."""

    def __init__(self, mol, aux):
        p = C.c_void_p()
        check(lib().hh_cint_pair_create(mol._h, aux._h, C.byref(p)))
        self._p = p
        self._keep = (mol, aux)
        self.fn = C.cast(lib().hh_cint_int3c_deriv, C.c_void_p)
        self.ctx = p

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                lib().hh_cint_pair_destroy(p)
            except Exception:
                pass
            self._p = None
