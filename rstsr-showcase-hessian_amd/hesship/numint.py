"""Grid contractions (reference: src/numint_matmul) and the XC kernel transform (src/xceff).

Shapes (reference, column-major): ao [ng, nao, ncomp]; weights [ng]; rho1 [ng, nvar, nprop];
fxc_eff [ng, nvar, nvar]; bra/mocc [nao, nocc]; kets / mo1_bra [nao, nocc, nprop].
den_type: "rho" | "sigma" | "tau" (XCDenType, src/xceff/flags.rs:19-24).
"""
import ctypes as C

import numpy as np

from ._lib import PREC, check, lib
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


def transform_xc(rho, xc_val, den_type, order):
    """xc_deriv.rs:305-428 (unpolarised): vxc_eff [ng, nvar] (order 1) / fxc_eff [ng, nvar, nvar] (order 2)."""
    ng = xc_val.shape[0]
    nv = NVAR[den_type]
    shape = (ng,) + (nv,) * order
    out = empty_f(shape, xc_val.device)
    r = fortran(rho) if rho is not None else None
    x = fortran(xc_val)
    check(lib().hh_xceff_transform(DEN[den_type], int(order), r.data_ptr() if r is not None else None,
                                   r.shape[0] if r is not None else 0, x.data_ptr(), x.shape[0], ng, out.data_ptr(),
                                   stream_of(xc_val.device)))
    return out
