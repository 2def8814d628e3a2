"""Block Krylov solver and dense helpers (reference: src/util/krylov_block.rs)."""
import ctypes as C

import numpy as np
import torch

from ._lib import AOP_FN, KrylovStats, check, lib
from .tensors import empty_f, fortran, stream_of


def krylov_block(aop, b, x0=None, tol=1e-8, max_cycle=42, max_space=14, lindep=1e-14, stats=None):
    """krylov_block.rs:39-47.  aop(x [n, nb] torch) -> [n, nb]; b [n, nset] device tensor.

    Raises hesship.KrylovNotConverged where the reference panics."""
    n, nset = b.shape
    dev = b.device
    bf = fortran(b)
    x0f = fortran(x0) if x0 is not None else None
    x = empty_f((n, nset), dev)
    st = KrylovStats()
    stream = stream_of(dev)
    err = []

    def cb(ctx, xp, nn, nb, axp):
        try:
            xt = empty_f((n, nb), dev)
            check(lib().hh_util_copy(xt.data_ptr(), xp, 8 * n * nb, stream))
            y = fortran(aop(xt))
            assert tuple(y.shape) == (n, nb), f"aop returned {tuple(y.shape)}, expected {(n, nb)}"
            check(lib().hh_util_copy(axp, y.data_ptr(), 8 * n * nb, stream))
            torch.cuda.current_stream(dev).synchronize()
            return 0
        except Exception as e:  # surfaced after the call returns
            err.append(e)
            return 1

    fn = AOP_FN(cb)
    rc = lib().hh_krylov_block(fn, None, bf.data_ptr(), x0f.data_ptr() if x0f is not None else None, n, nset,
                               float(tol), int(max_cycle), int(max_space), float(lindep), x.data_ptr(),
                               C.byref(st), stream)
    if stats is not None:
        stats.update(st.as_dict())
    if err:
        raise err[0]
    check(rc)
    return x


def orth_block(vec, lindep=1e-14):
    """orth_block (krylov_block.rs:225-253) on the device: non-normalised Gram-Schmidt of the
    columns of vec [n, m] in order, dropping columns with ||v||^2 <= lindep.  Returns (kept
    columns [n, kept], list of their squared norms)."""
    n, m = vec.shape
    V = fortran(vec).clone()
    R = empty_f((n, max(m, 1)), vec.device)
    nr = np.zeros(max(m, 1))
    k = C.c_int64()
    check(lib().hh_krylov_orth_block(V.data_ptr(), n, m, float(lindep), R.data_ptr(), nr.ctypes.data, C.byref(k),
                                     stream_of(vec.device)))
    return R[:, :k.value], list(nr[:k.value])


def solve(h, g):
    """Dense FP64 solve h x = g on the device (copies; returns x)."""
    H = fortran(h).clone()
    G = fortran(g.reshape(g.shape[0], -1)).clone()
    check(lib().hh_linalg_solve(H.data_ptr(), H.shape[0], G.data_ptr(), G.shape[1], stream_of(h.device)))
    return G.reshape(g.shape)


def eigh(a):
    """Symmetric eigendecomposition on the device (Jacobi): (w ascending, V columns); reads the
    lower triangle of a."""
    V = fortran(a).clone()
    w = torch.empty(V.shape[0], dtype=torch.float64, device=a.device)
    check(lib().hh_linalg_eigh(V.data_ptr(), V.shape[0], w.data_ptr(), stream_of(a.device)))
    return w, V


def cholesky(a):
    L = fortran(a).clone()
    check(lib().hh_linalg_cholesky(L.data_ptr(), L.shape[0], stream_of(a.device)))
    return L


def trsm_lower(L, x, side_right=False):
    """side_right False: L^-1 x (x [nd, n2]);  True: x L^-T (x [n2, nd])."""
    X = fortran(x).clone()
    Lf = fortran(L)  # held across the call
    nd = L.shape[0]
    n2 = X.shape[0] if side_right else X.shape[1]
    check(lib().hh_linalg_trsm_lower(Lf.data_ptr(), nd, X.data_ptr(), n2, X.shape[0], int(side_right),
                                     stream_of(x.device)))
    return X
