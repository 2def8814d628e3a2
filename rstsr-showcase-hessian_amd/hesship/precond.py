"""Learned diagonal CPHF preconditioner (no reference counterpart).

The reference's dimensionless CPHF (rscf.rs:237-263) is preconditioned only by the orbital-energy
scaling.  Here a small MLP (3 -> 64 -> 64 -> 64 -> 1, ReLU, FP32; trained and evaluated by the
library's own kernels) maps per-(a,i) features -- ln(e_a - e_i), ln of the RI pair norm (ai|ai),
ln of the current residual magnitude -- to ln D[a,i], a right diagonal preconditioner
(x = D y).  Training pairs (residual, solution) come from the method's own FP64 solves.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import check, lib
from .tensors import empty_f, fortran, stream_of

NPARAM = 8641


class DiagMLP:
    def __init__(self, seed=0x5EED0005, device="cuda:0"):
        self.device = torch.device(device)
        h = C.c_void_p()
        check(lib().hh_mlp_create(int(seed), stream_of(self.device), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().hh_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    def params(self):
        out = np.empty(NPARAM, dtype=np.float32)
        check(lib().hh_mlp_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().hh_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, x):
        xf = fortran(x, torch.float32)
        y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        check(lib().hh_mlp_forward(self._h, xf.data_ptr(), x.shape[0], y.data_ptr()))
        return y

    def loss_and_grad(self, x, t, w):
        xf, tf, wf = fortran(x, torch.float32), t.float().contiguous(), w.float().contiguous()
        g = torch.empty(NPARAM, dtype=torch.float32, device=x.device)
        loss = C.c_double()
        check(lib().hh_mlp_loss_and_grad(self._h, xf.data_ptr(), tf.data_ptr(), wf.data_ptr(), x.shape[0],
                                         g.data_ptr(), C.byref(loss)))
        return loss.value, g

    def train(self, x, t, w, epochs=300, lr=3e-3):
        xf, tf, wf = fortran(x, torch.float32), t.float().contiguous(), w.float().contiguous()
        loss = C.c_double()
        check(lib().hh_mlp_train(self._h, xf.data_ptr(), tf.data_ptr(), wf.data_ptr(), x.shape[0], int(epochs),
                                 float(lr), C.byref(loss)))
        return loss.value


def pair_norms(rijk, mo_coeff):
    c = fortran(mo_coeff)
    out = empty_f((c.shape[1], rijk.nocc[0]), c.device)
    check(lib().hh_rijk_pair_norms(rijk._h, 0, c.data_ptr(), c.shape[1], out.data_ptr()))
    return out


def features(drv, pair, r):
    nvir = drv.nmo - drv.nocc
    rr = fortran(r.reshape(drv.nmo, drv.nocc, -1))
    pr = fortran(pair)
    x = empty_f((nvir * drv.nocc, 3), r.device, torch.float32)
    check(lib().hh_cphf_precond_features(drv._h, pr.data_ptr(), rr.data_ptr(), rr.shape[2], x.data_ptr()))
    return x


def target(drv, r, xs):
    nvir = drv.nmo - drv.nocc
    rr = fortran(r.reshape(drv.nmo, drv.nocc, -1))
    xx = fortran(xs.reshape(drv.nmo, drv.nocc, -1))
    t = torch.empty(nvir * drv.nocc, dtype=torch.float32, device=r.device)
    w = torch.empty_like(t)
    check(lib().hh_cphf_precond_target(drv._h, rr.data_ptr(), xx.data_ptr(), rr.shape[2], t.data_ptr(), w.data_ptr()))
    return t, w


def diag(drv, y):
    yc = y.float().contiguous()
    d = empty_f((drv.nmo, drv.nocc), y.device)
    check(lib().hh_cphf_precond_diag(drv._h, yc.data_ptr(), d.data_ptr()))
    return d
