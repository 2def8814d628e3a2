"""Dimensionless CPHF drivers for restricted and unrestricted SCF (reference:
src/hessian/rscf.rs, src/hessian/uscf.rs)."""
import ctypes as C

import numpy as np
import torch

from ._lib import PREC, KrylovStats, check, lib, ptr_array
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

    def set_graphs(self, on=True):
        """Replay the response operator from HIP graphs (default on)."""
        check(lib().hh_cphf_set_graphs(self._h, int(bool(on))))

    def graph_stats(self):
        """(graphs captured, graph launches) so far."""
        c, l = C.c_int64(), C.c_int64()
        check(lib().hh_cphf_graph_stats(self._h, C.byref(c), C.byref(l)))
        return c.value, l.value

    def finalize_reuses(self):
        """finalize calls that reused the FP64 response from the last mixed-precision solve."""
        c = C.c_int64()
        check(lib().hh_cphf_finalize_reuses(self._h, C.byref(c)))
        return c.value

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


def _host(x):
    return np.ascontiguousarray(np.asarray(x.cpu() if isinstance(x, torch.Tensor) else x, dtype=np.float64))


class UHessCPHF:
    """The CPHF part of UHessSCF (uscf.rs:32-446) over device tensors: every argument and result
    is a pair (alpha, beta).  mo_coeff[s] [nao, nmo] (device); mo_occ[s], mo_energy[s] [nmo]."""

    def __init__(self, mo_coeff, mo_occ, mo_energy, config=None):
        cfg = dict(DEFAULT_CONFIG, **(config or {}))
        self.device = mo_coeff[0].device
        self.nao = mo_coeff[0].shape[0]
        self.nmo = [int(c.shape[1]) for c in mo_coeff]
        occ = [_host(o) for o in mo_occ]
        en = [_host(e) for e in mo_energy]
        self.nocc = [int((o > 0).sum()) for o in occ]
        self._c = [fortran(c) for c in mo_coeff]
        self._host = (occ, en)
        nmo = (C.c_int64 * 2)(*self.nmo)
        h = C.c_void_p()
        check(lib().hh_ucphf_create(ptr_array([c.data_ptr() for c in self._c]), self.nao, nmo,
                                    ptr_array([o.ctypes.data for o in occ]), ptr_array([e.ctypes.data for e in en]),
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

    def set_graphs(self, on=True):
        """Replay the response operator from HIP graphs (default on)."""
        check(lib().hh_cphf_set_graphs(self._h, int(bool(on))))

    def graph_stats(self):
        """(graphs captured, graph launches) so far."""
        c, l = C.c_int64(), C.c_int64()
        check(lib().hh_cphf_graph_stats(self._h, C.byref(c), C.byref(l)))
        return c.value, l.value

    def finalize_reuses(self):
        """finalize calls that reused the FP64 response from the last mixed-precision solve."""
        c = C.c_int64()
        check(lib().hh_cphf_finalize_reuses(self._h, C.byref(c)))
        return c.value

    def add_el(self, rijk):
        """Append a UHF RI-JK component (UHessRIJK, prepared) to the response list."""
        check(lib().hh_cphf_add_rijk(self._h, rijk._h))
        self._els.append(rijk)

    def add_numint(self, ni, den_type, fxc_eff):
        """Append the UKS grid response (UHessKSNIMatmul, hess_uks.rs:740-770) with fxc_eff
        [ng, nvar, 2, nvar, 2]."""
        from .numint import DEN
        f = fortran(fxc_eff)
        check(lib().hh_ucphf_add_numint(self._h, ni._h, DEN[den_type], f.data_ptr()))
        self._els.append((ni, f))

    def _flat(self, pair):
        shapes = [tuple(t.shape) for t in pair]
        return [fortran(t.reshape(t.shape[0], t.shape[1], -1)) for t in pair], shapes

    def _out(self, like):
        return [empty_f(tuple(t.shape), self.device) for t in like]

    def compute_dimless_rhs(self, f1mo, s1mo):
        f, shapes = self._flat(f1mo)
        s, _ = self._flat(s1mo)
        out = self._out(f)
        check(lib().hh_ucphf_dimless_rhs(self._h, ptr_array([x.data_ptr() for x in f]),
                                         ptr_array([x.data_ptr() for x in s]), f[0].shape[2],
                                         ptr_array([x.data_ptr() for x in out])))
        return [o.reshape(sh) for o, sh in zip(out, shapes)]

    def response_mo(self, mo1, precision="fp64"):
        m, shapes = self._flat(mo1)
        out = self._out(m)
        check(lib().hh_ucphf_response_mo(self._h, ptr_array([x.data_ptr() for x in m]), m[0].shape[2],
                                         ptr_array([x.data_ptr() for x in out]), PREC[precision]))
        return [o.reshape(sh) for o, sh in zip(out, shapes)]

    def response_dimless_cphf(self, mo1, precision="fp64"):
        m, shapes = self._flat(mo1)
        out = self._out(m)
        check(lib().hh_ucphf_response_dimless(self._h, ptr_array([x.data_ptr() for x in m]), m[0].shape[2],
                                              ptr_array([x.data_ptr() for x in out]), PREC[precision]))
        return [o.reshape(sh) for o, sh in zip(out, shapes)]

    def solve_dimless_cphf(self, rhs, precision="fp64", stats=None):
        r, shapes = self._flat(rhs)
        out = self._out(r)
        st = KrylovStats()
        rc = lib().hh_ucphf_solve_dimless(self._h, ptr_array([x.data_ptr() for x in r]), r[0].shape[2],
                                          ptr_array([x.data_ptr() for x in out]), PREC[precision], C.byref(st))
        if stats is not None:
            stats.update(st.as_dict())
        check(rc)
        return [o.reshape(sh) for o, sh in zip(out, shapes)]

    def finalize_cphf(self, f1mo, s1mo, mo1):
        f, shapes = self._flat(f1mo)
        s, _ = self._flat(s1mo)
        m, _ = self._flat(mo1)
        nprop = f[0].shape[2]
        mo1_out = self._out(f)
        mo_e1 = [empty_f((no, no, nprop), self.device) for no in self.nocc]
        check(lib().hh_ucphf_finalize(self._h, ptr_array([x.data_ptr() for x in f]),
                                      ptr_array([x.data_ptr() for x in s]), ptr_array([x.data_ptr() for x in m]), nprop,
                                      ptr_array([x.data_ptr() for x in mo1_out]),
                                      ptr_array([x.data_ptr() for x in mo_e1])))
        return ([o.reshape(sh) for o, sh in zip(mo1_out, shapes)],
                [e.reshape((no, no) + sh[2:]) for e, no, sh in zip(mo_e1, self.nocc, shapes)])

    def get_cphf_hess(self, f1mo, s1mo, mo1, mo_e1):
        natm = f1mo[0].shape[-1]
        f = [fortran(x.reshape(nm, no, 3, natm)) for x, nm, no in zip(f1mo, self.nmo, self.nocc)]
        s = [fortran(x.reshape(nm, no, 3, natm)) for x, nm, no in zip(s1mo, self.nmo, self.nocc)]
        m = [fortran(x.reshape(nm, no, 3, natm)) for x, nm, no in zip(mo1, self.nmo, self.nocc)]
        e = [fortran(x.reshape(no, no, 3, natm)) for x, no in zip(mo_e1, self.nocc)]
        de = empty_f((3, 3, natm, natm), self.device)
        check(lib().hh_ucphf_hess(self._h, ptr_array([x.data_ptr() for x in f]), ptr_array([x.data_ptr() for x in s]),
                                  ptr_array([x.data_ptr() for x in m]), ptr_array([x.data_ptr() for x in e]), natm,
                                  de.data_ptr()))
        return de

    def make_cphf_hess(self, f1mo, s1mo, precision="fp64", stats=None):
        """uscf.rs:442-446: rhs -> Krylov -> finalize -> de_cphf."""
        rhs = self.compute_dimless_rhs(f1mo, s1mo)
        mo1 = self.solve_dimless_cphf(rhs, precision=precision, stats=stats)
        mo1, mo_e1 = self.finalize_cphf(f1mo, s1mo, mo1)
        return self.get_cphf_hess(f1mo, s1mo, mo1, mo_e1)


class CPHFBatch:
    """Many dimensionless CPHF solves advanced in lockstep (no reference counterpart: the
    reference runs one RHessSCF / UHessSCF per molecule, rscf.rs:481-505).  Each system is a
    driver (RHessCPHF or UHessCPHF) plus the RI-JK handle prepared with its orbitals, whose
    resident response matrix is the system's operator; each is solved by the reference's
    krylov_block algorithm in FP64 with one subspace (no restart), every Krylov step one grouped
    launch over all running systems."""

    def __init__(self, tol=1e-8, max_cycle=42, lindep=1e-14, device=None):
        dev = device or torch.device("cuda", torch.cuda.current_device())
        h = C.c_void_p()
        check(lib().hh_cphf_batch_create(tol, max_cycle, lindep, stream_of(dev), C.byref(h)))
        self._h = h
        self._keep = []
        self._out = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().hh_cphf_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def add(self, drv, rijk, rhs):
        """Add a system: rhs [nmo, nocc, ...] (restricted) or a pair of them (unrestricted).
        Returns the index of the system."""
        pair = isinstance(drv, UHessCPHF)
        rl = list(rhs) if pair else [rhs]
        flat = [fortran(t.reshape(t.shape[0], t.shape[1], -1)) for t in rl]
        nprop = flat[0].shape[2]
        out = [empty_f(tuple(t.shape), t.device) for t in flat]
        check(lib().hh_cphf_batch_add(self._h, drv._h, rijk._h, ptr_array([x.data_ptr() for x in flat]), nprop,
                                      ptr_array([x.data_ptr() for x in out])))
        self._keep.append((drv, rijk, flat))
        self._out.append((pair, [o.reshape(t.shape) for o, t in zip(out, rl)]))
        return len(self._out) - 1

    def solve(self, stats=None):
        """Solve every system; returns the mo1 of each (a tensor, or a pair for unrestricted).
        stats (list, optional) receives one dict per system."""
        n = len(self._out)
        st = (KrylovStats * max(n, 1))()
        rc = lib().hh_cphf_batch_solve(self._h, C.cast(st, C.c_void_p))
        if stats is not None:
            stats.extend(st[i].as_dict() for i in range(n))
        check(rc)
        return [o if pair else o[0] for pair, o in self._out]

    def info(self):
        """(systems, lockstep cycles of the last solve, most launches in one cycle)."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().hh_cphf_batch_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def times(self):
        """Wall seconds of the last solve: matrix builds, set-up and first block, lockstep cycles,
        projected solves."""
        t = (C.c_double * 4)()
        check(lib().hh_cphf_batch_times(self._h, t))
        return tuple(t)
