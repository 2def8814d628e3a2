"""Dimensionless CPHF driver for restricted SCF (reference: src/hessian/rscf.rs)."""
import ctypes as C

import numpy as np
import torch

from ._lib import PREC, KrylovStats, check, lib
from .tensors import empty_f, fortran, stream_of

DEFAULT_CONFIG = dict(level_shift=0.0, cphf_tol=1e-8, cphf_max_cycle=42, cphf_max_space=14, cphf_lindep=1e-14)


class RHessCPHF:
    """The CPHF part of RHessSCF (rscf.rs:21-425) over device tensors.

    mo_coeff [nao, nmo] (device); mo_occ, mo_energy [nmo] (host or device, copied to host)."""

    def __init__(self, mo_coeff, mo_occ, mo_energy, config=None):
        cfg = dict(DEFAULT_CONFIG, **(config or {}))
        self.device = mo_coeff.device
        self.nao, self.nmo = mo_coeff.shape
        occ = np.ascontiguousarray(np.asarray(mo_occ.cpu() if isinstance(mo_occ, torch.Tensor) else mo_occ, dtype=np.float64))
        en = np.ascontiguousarray(np.asarray(mo_energy.cpu() if isinstance(mo_energy, torch.Tensor) else mo_energy, dtype=np.float64))
        self.nocc = int((occ > 0).sum())
        self._c = fortran(mo_coeff)
        h = C.c_void_p()
        check(lib().hh_cphf_create(self._c.data_ptr(), self.nao, self.nmo, occ.ctypes.data, en.ctypes.data,
                                   cfg["level_shift"], cfg["cphf_tol"], cfg["cphf_max_cycle"], cfg["cphf_max_space"],
                                   cfg["cphf_lindep"], stream_of(self.device), C.byref(h)))
        self._h = h
        self._els = []
        self.config = cfg

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().hh_cphf_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_verbose(self, on=True):
        check(lib().hh_cphf_set_verbose(self._h, int(bool(on))))

    def add_el(self, rijk):
        """Append an RHF RI-JK component (prepared) to the response list."""
        check(lib().hh_cphf_add_rijk(self._h, rijk._h))
        self._els.append(rijk)

    def add_numint(self, ni, den_type, fxc_eff):
        """Append the RKS grid response (RHessKSNIMatmul, hess_rks.rs:803-909) with fxc_eff."""
        from .numint import DEN
        f = fortran(fxc_eff)
        check(lib().hh_cphf_add_numint(self._h, ni._h, DEN[den_type], f.data_ptr()))
        self._els.append((ni, f))

    def _flat(self, t):
        shape = tuple(t.shape)
        return fortran(t.reshape(shape[0], shape[1], -1)), shape

    def compute_dimless_rhs(self, f1mo, s1mo):
        f, shape = self._flat(f1mo)
        s, _ = self._flat(s1mo)
        out = empty_f(tuple(f.shape), self.device)
        check(lib().hh_cphf_dimless_rhs(self._h, f.data_ptr(), s.data_ptr(), f.shape[2], out.data_ptr()))
        return out.reshape(shape)

    def response_mo(self, mo1, precision="fp64"):
        m, shape = self._flat(mo1)
        out = empty_f(tuple(m.shape), self.device)
        check(lib().hh_cphf_response_mo(self._h, m.data_ptr(), m.shape[2], out.data_ptr(), PREC[precision]))
        return out.reshape(shape)

    def response_dimless_cphf(self, mo1, precision="fp64"):
        m, shape = self._flat(mo1)
        out = empty_f(tuple(m.shape), self.device)
        check(lib().hh_cphf_response_dimless(self._h, m.data_ptr(), m.shape[2], out.data_ptr(), PREC[precision]))
        return out.reshape(shape)

    def solve_dimless_cphf(self, rhs, precision="fp64", diag=None, stats=None):
        r, shape = self._flat(rhs)
        out = empty_f(tuple(r.shape), self.device)
        st = KrylovStats()
        d = fortran(diag.reshape(self.nmo, self.nocc)) if diag is not None else None
        rc = lib().hh_cphf_solve_dimless(self._h, r.data_ptr(), r.shape[2], out.data_ptr(), PREC[precision],
                                         d.data_ptr() if d is not None else None, C.byref(st))
        if stats is not None:
            stats.update(st.as_dict())
        check(rc)
        return out.reshape(shape)

    def finalize_cphf(self, f1mo, s1mo, mo1):
        f, shape = self._flat(f1mo)
        s, _ = self._flat(s1mo)
        m, _ = self._flat(mo1)
        nprop = f.shape[2]
        mo1_out = empty_f(tuple(f.shape), self.device)
        mo_e1 = empty_f((self.nocc, self.nocc, nprop), self.device)
        check(lib().hh_cphf_finalize(self._h, f.data_ptr(), s.data_ptr(), m.data_ptr(), nprop, mo1_out.data_ptr(),
                                     mo_e1.data_ptr()))
        return mo1_out.reshape(shape), mo_e1.reshape((self.nocc, self.nocc) + shape[2:])

    def get_cphf_hess(self, f1mo, s1mo, mo1, mo_e1):
        natm = f1mo.shape[-1]
        f = fortran(f1mo.reshape(self.nmo, self.nocc, 3, natm))
        s = fortran(s1mo.reshape(self.nmo, self.nocc, 3, natm))
        m = fortran(mo1.reshape(self.nmo, self.nocc, 3, natm))
        e = fortran(mo_e1.reshape(self.nocc, self.nocc, 3, natm))
        de = empty_f((3, 3, natm, natm), self.device)
        check(lib().hh_cphf_hess(self._h, f.data_ptr(), s.data_ptr(), m.data_ptr(), e.data_ptr(), natm, de.data_ptr()))
        return de

    def make_cphf_hess(self, f1mo, s1mo, precision="fp64", stats=None):
        """rscf.rs:412-425: rhs -> Krylov -> finalize -> de_cphf."""
        rhs = self.compute_dimless_rhs(f1mo, s1mo)
        mo1 = self.solve_dimless_cphf(rhs, precision=precision, stats=stats)
        mo1, mo_e1 = self.finalize_cphf(f1mo, s1mo, mo1)
        return self.get_cphf_hess(f1mo, s1mo, mo1, mo_e1)
