"""Grid contractions (reference: src/numint_matmul) and the XC kernel transform (src/xceff).

Shapes (reference, column-major): ao [ng, nao, ncomp]; weights [ng]; rho1 [ng, nvar, nprop];
fxc_eff [ng, nvar, nvar]; bra/mocc [nao, nocc]; kets / mo1_bra [nao, nocc, nprop].  UKS: pairs
(alpha, beta) of the per-spin arrays; rho1 [ng, nvar, 2, nprop]; fxc_eff [ng, nvar, 2, nvar, 2].
den_type: "rho" | "sigma" | "tau" (XCDenType, src/xceff/flags.rs:19-24).
"""
import ctypes as C

import numpy as np
import torch

from ._lib import PREC, check, lib, ptr_array
from .tensors import empty_f, fortran, stream_of

DEN = {"rho": 0, "sigma": 1, "tau": 2}
NVAR = {"rho": 1, "sigma": 4, "tau": 5}


class NIMatmul:
    """Grid handle over precomputed AO values (NIMatmul, nimatmul.rs:8-37)."""

    def __init__(self, ao, weights, nbatch=0):
        self.device = ao.device
        self.ng, self.nao, self.ncomp = ao.shape
        self._ao = fortran(ao)
        self._w = fortran(weights)
        h = C.c_void_p()
        check(lib().hh_numint_create(self._ao.data_ptr(), self.ng, self.nao, self.ncomp, self._w.data_ptr(),
                                     stream_of(self.device), C.byref(h)))
        self._h = h
        if nbatch:
            check(lib().hh_numint_set_batch(self._h, int(nbatch)))

    def set_fused(self, on=True):
        """BF16X3 RKS response: kets product fused with the pointwise pass (default on)."""
        check(lib().hh_numint_set_fused(self._h, int(bool(on))))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().hh_numint_destroy(h)
            except Exception:
                pass
            self._h = None

    # Every device operand is bound to a local name for the duration of the C call: a temporary
    # copy made inline in the argument list would be freed (and its block reused) before the
    # asynchronous kernels read it.
    def make_rho_from_one_bra_mult_ket(self, bra, kets, den_type, precision="fp64"):
        nao, nocc = bra.shape
        k = fortran(kets.reshape(nao, nocc, -1))
        b = fortran(bra)
        nprop = k.shape[2]
        out = empty_f((self.ng, NVAR[den_type], nprop), self.device)
        check(lib().hh_numint_rho1_one_bra_mult_ket(self._h, DEN[den_type], b.data_ptr(), nocc,
                                                    k.data_ptr(), nprop, out.data_ptr(), PREC[precision]))
        return out

    def make_rks_fxc_pot_with_eff_bra_trans(self, fxc_eff, rho1, bra, den_type, precision="fp64"):
        nao, nocc = bra.shape
        nprop = rho1.shape[2]
        f, r, b = fortran(fxc_eff), fortran(rho1), fortran(bra)
        out = empty_f((nao, nocc, nprop), self.device)
        check(lib().hh_numint_fxc_pot_bra_trans(self._h, DEN[den_type], f.data_ptr(), r.data_ptr(), nprop,
                                                b.data_ptr(), nocc, out.data_ptr(), PREC[precision]))
        return out

    def get_rks_response_bra(self, den_type, fxc_eff, mo1_bra, mocc, precision="fp64"):
        """hess_rks.rs:570-605 (x4 closed-shell factor), batched over the grid."""
        shape = tuple(mo1_bra.shape)
        nao, nocc = mocc.shape
        k = fortran(mo1_bra.reshape(nao, nocc, -1))
        f, m = fortran(fxc_eff), fortran(mocc)
        nprop = k.shape[2]
        out = empty_f((nao, nocc, nprop), self.device)
        check(lib().hh_numint_rks_response_bra(self._h, DEN[den_type], f.data_ptr(), k.data_ptr(), nprop,
                                               m.data_ptr(), nocc, out.data_ptr(), PREC[precision]))
        return out.reshape(shape)

    def make_uks_fxc_pot_with_eff_bra_trans(self, fxc_eff, rho1, bra, den_type, precision="fp64"):
        """pure_xcpot.rs:921-1028: [out_a, out_b], out_s [nao, nocc_s, nprop]."""
        nprop = rho1.shape[3]
        f, r = fortran(fxc_eff), fortran(rho1)
        b = [fortran(x) for x in bra]
        nocc = (C.c_int64 * 2)(*[int(x.shape[1]) for x in b])
        out = [empty_f((self.nao, int(x.shape[1]), nprop), self.device) for x in b]
        check(lib().hh_numint_uks_fxc_pot_bra_trans(self._h, DEN[den_type], f.data_ptr(), r.data_ptr(), nprop,
                                                    ptr_array([x.data_ptr() for x in b]), nocc,
                                                    ptr_array([o.data_ptr() for o in out]), PREC[precision]))
        return out

    def get_uks_response_bra(self, den_type, fxc_eff, mo1_bra, mocc, precision="fp64"):
        """hess_uks.rs:390-441 (x2 factor), batched over the grid: [resp_a, resp_b]."""
        m = [fortran(x) for x in mocc]
        nocc_l = [int(x.shape[1]) for x in m]
        shapes = [tuple(x.shape) for x in mo1_bra]
        k = [fortran(x.reshape(self.nao, n, -1)) for x, n in zip(mo1_bra, nocc_l)]
        nprop = k[0].shape[2]
        if k[1].shape[2] != nprop:
            raise ValueError("the two spins' mo1_bra differ in the number of perturbations")
        f = fortran(fxc_eff)
        out = [empty_f((self.nao, n, nprop), self.device) for n in nocc_l]
        check(lib().hh_numint_uks_response_bra(self._h, DEN[den_type], f.data_ptr(),
                                               ptr_array([x.data_ptr() for x in k]), nprop,
                                               ptr_array([x.data_ptr() for x in m]), (C.c_int64 * 2)(*nocc_l),
                                               ptr_array([o.data_ptr() for o in out]), PREC[precision]))
        return [o.reshape(s) for o, s in zip(out, shapes)]

    # ---- DFT skeleton-Hessian setup (hess_rks.rs:71-750, hess_uks.rs:56-384) ----
    @staticmethod
    def _spins(dm0):
        return list(dm0) if isinstance(dm0, (list, tuple)) else [dm0]

    def vxc_matrix(self, den_type, vxc_eff):
        """The exchange-correlation potential matrix [nao, nao] = dE_xc / dD from vxc_eff [ng, nvar]."""
        v = fortran(vxc_eff)
        out = empty_f((self.nao, self.nao), self.device)
        check(lib().hh_numint_vxc_matrix(self._h, DEN[den_type], v.data_ptr(), out.data_ptr()))
        return out

    def setup_rho(self, den_type, dm0):
        """rho [ng, nvar] (RKS, dm0 a matrix) or [ng, nvar, 2] (UKS, dm0 a pair)."""
        d = [fortran(x) for x in self._spins(dm0)]
        ns = len(d)
        out = empty_f((self.ng, NVAR[den_type], ns), self.device)
        check(lib().hh_numint_setup_rho(self._h, DEN[den_type], ns, ptr_array([x.data_ptr() for x in d]),
                                        out.data_ptr()))
        return out if ns == 2 else out[:, :, 0]

    def make_hessian_setup(self, den_type, dm0, aoslices, vxc_eff, fxc_eff):
        """The additive intermediates of make_hessian_setup_batched with the reference's keys
        (UKS: per-spin keys with suffix _a / _b).  aoslices: (p0, p1) per listed atom."""
        d = [fortran(x) for x in self._spins(dm0)]
        ns = len(d)
        natm = len(aoslices)
        sl = (C.c_int64 * (2 * natm))(*[int(v) for p in aoslices for v in p])
        v, f = fortran(vxc_eff), fortran(fxc_eff)
        nao, dev = self.nao, self.device
        de_fxc = empty_f((3, 3, natm, natm), dev)
        per = {k: [empty_f(shape, dev) for _ in range(ns)] for k, shape in
               [("de_vxc_diag", (3, 3, natm, natm)), ("de_vxc_off", (3, 3, natm, natm)), ("vmat_ip", (nao, nao, 3)),
                ("vmat_deriv1", (nao, nao, 3, natm))]}
        pa = lambda k: ptr_array([t.data_ptr() for t in per[k]])
        check(lib().hh_numint_hessian_setup(self._h, DEN[den_type], ns, ptr_array([x.data_ptr() for x in d]), natm, sl,
                                            v.data_ptr(), f.data_ptr(), de_fxc.data_ptr(), pa("de_vxc_diag"),
                                            pa("de_vxc_off"), pa("vmat_ip"), pa("vmat_deriv1")))
        out = {"de_fxc": de_fxc}
        for k, ts in per.items():
            if ns == 1:
                out[k] = ts[0]
            else:
                out[k + "_a"], out[k + "_b"] = ts
        return out

    def make_hessian_setup_becke(self, den_type, dm0, aoslices, exc, vxc_eff, fxc_eff, becke, grid_coords,
                                 quadrature_weights, atm_quad_split, hardness=3, grid_fixed_keys=True):
        """make_hessian_setup_becke (hess_rks_becke.rs:1106-1266; UKS hess_uks_becke.rs:726-930):
        the grid is grouped by generating atom (atm_quad_split [natm + 1]) and moves with the atoms;
        the handle's weights are the partitioned ones, quadrature_weights the ones before the
        partition.  exc [ng] is the energy per particle.  Returns de_xc_skeleton and
        vmat_deriv1_grid (UKS: _a / _b) and, with grid_fixed_keys, the keys of
        make_hessian_setup and de_xc_skeleton_no_becke.  becke: BeckeMolTables of all atoms."""
        d = [fortran(x) for x in self._spins(dm0)]
        ns = len(d)
        natm = len(aoslices)
        sl = (C.c_int64 * (2 * natm))(*[int(v) for p in aoslices for v in p])
        split = np.ascontiguousarray(np.asarray(atm_quad_split, dtype=np.int64))
        if split.shape != (natm + 1,):
            raise ValueError("atm_quad_split must have length natm + 1")
        v, f, e = fortran(vxc_eff), fortran(fxc_eff), fortran(exc)
        coords = grid_coords.to(device=self.device, dtype=torch.float64).contiguous()
        wq = fortran(quadrature_weights)
        nao, dev = self.nao, self.device
        de = empty_f((3, 3, natm, natm), dev)
        grid = [empty_f((nao, nao, 3, natm), dev) for _ in range(ns)]
        per = {}
        de_fxc = None
        if grid_fixed_keys:
            de_fxc = empty_f((3, 3, natm, natm), dev)
            per = {k: [empty_f(shape, dev) for _ in range(ns)] for k, shape in
                   [("de_vxc_diag", (3, 3, natm, natm)), ("de_vxc_off", (3, 3, natm, natm)),
                    ("vmat_ip", (nao, nao, 3)), ("vmat_deriv1", (nao, nao, 3, natm))]}
        pa = lambda k: ptr_array([t.data_ptr() for t in per[k]]) if k in per else None
        check(lib().hh_numint_hessian_setup_becke(
            self._h, DEN[den_type], ns, ptr_array([x.data_ptr() for x in d]), natm, sl, e.data_ptr(), v.data_ptr(),
            f.data_ptr(), becke._h, coords.data_ptr(), wq.data_ptr(), split.ctypes.data_as(C.POINTER(C.c_int64)),
            int(hardness), de.data_ptr(), ptr_array([g.data_ptr() for g in grid]),
            de_fxc.data_ptr() if de_fxc is not None else None, pa("de_vxc_diag"), pa("de_vxc_off"), pa("vmat_ip"),
            pa("vmat_deriv1")))
        out = {"de_xc_skeleton": de}
        if ns == 1:
            out["vmat_deriv1_grid"] = grid[0]
        else:
            out["vmat_deriv1_grid_a"], out["vmat_deriv1_grid_b"] = grid
        if grid_fixed_keys:
            out["de_fxc"] = de_fxc
            for k, ts in per.items():
                if ns == 1:
                    out[k] = ts[0]
                else:
                    out[k + "_a"], out[k + "_b"] = ts
            out["de_xc_skeleton_no_becke"] = self.skeleton_hess(out)
        return out

    @staticmethod
    def skeleton_hess(setup):
        """make_skeleton_hess (hess_rks.rs:866-872 / hess_uks.rs:705-720)."""
        tags = [""] if "de_vxc_diag" in setup else ["_a", "_b"]
        return setup["de_fxc"] + sum(setup["de_vxc_diag" + t] + setup["de_vxc_off" + t] for t in tags)


def _check_xc_columns(xc_val, xlen, order):
    if order not in (1, 2):
        raise ValueError("order must be 1 (vxc) or 2 (fxc)")
    need = 1 + xlen + (xlen * (xlen + 1) // 2 if order == 2 else 0)
    if xc_val.dim() != 2 or xc_val.shape[1] < need:
        raise ValueError(f"xc_val needs [ng, >= {need}] columns for order {order}")


def transform_xc(rho, xc_val, den_type, order):
    """xc_deriv.rs:305-428 (unpolarised): vxc_eff [ng, nvar] (order 1) / fxc_eff [ng, nvar, nvar] (order 2)."""
    ng = xc_val.shape[0]
    nv = NVAR[den_type]
    _check_xc_columns(xc_val, {"rho": 1, "sigma": 2, "tau": 3}[den_type], order)
    shape = (ng,) + (nv,) * order
    out = empty_f(shape, xc_val.device)
    r = fortran(rho) if rho is not None else None
    x = fortran(xc_val)
    check(lib().hh_xceff_transform(DEN[den_type], int(order), r.data_ptr() if r is not None else None,
                                   r.shape[0] if r is not None else 0, x.data_ptr(), x.shape[0], ng, out.data_ptr(),
                                   stream_of(xc_val.device)))
    return out


def transform_xc_polarized(rho, xc_val, den_type, order):
    """transform_xc_inner with LibXCSpin::Polarized (xc_deriv.rs:188-243, 305-428): rho
    [ng, >= nvar, 2]; xc_val [ng, ncol] in XCFun order; vxc_eff [ng, nvar, 2] (order 1) /
    fxc_eff [ng, nvar, 2, nvar, 2] (order 2)."""
    ng = xc_val.shape[0]
    nv = NVAR[den_type]
    _check_xc_columns(xc_val, {"rho": 2, "sigma": 5, "tau": 7}[den_type], order)
    if rho is not None and (rho.dim() != 3 or rho.shape[1] < nv or rho.shape[2] != 2 or rho.shape[0] != ng):
        raise ValueError(f"rho must be [ng, >= {nv}, 2]")
    shape = (ng,) + (nv, 2) * order
    out = empty_f(shape, xc_val.device)
    r = fortran(rho) if rho is not None else None
    x = fortran(xc_val)
    ld_rho = r.shape[0] if r is not None else 0
    ld_spin = r.shape[0] * r.shape[1] if r is not None else 0
    check(lib().hh_xceff_transform_polarized(DEN[den_type], int(order), r.data_ptr() if r is not None else None,
                                             ld_rho, ld_spin, x.data_ptr(), x.shape[0], ng, out.data_ptr(),
                                             stream_of(xc_val.device)))
    return out


def libxc_to_xcfun_mapping(den_type, deriv):
    """xc_deriv.rs:65-72 (spin polarised): the LibXC output columns listed in XCFun order."""
    n = C.c_int64()
    check(lib().hh_xceff_libxc_to_xcfun_map(DEN[den_type], int(deriv), None, 0, C.byref(n)))
    buf = (C.c_int64 * n.value)()
    check(lib().hh_xceff_libxc_to_xcfun_map(DEN[den_type], int(deriv), buf, n.value, C.byref(n)))
    return np.array(buf[:], dtype=np.int64)
