"""RI-JK response components (reference: src/ri_jk/hess_r.rs, src/ri_jk/hess_u.rs).

Shapes (reference, column-major):  cderi [nao_tp, naux];  mo_coeff [nao, nmo];  mo_occ [nmo];
bra [nao, nocc, ...] (trailing dims are the perturbations, collectively nprop).
"""
import ctypes as C

import numpy as np
import torch

from ._lib import PREC, check, lib, ptr_array
from .tensors import empty_f, fortran, stream_of

# RI metric decomposition policies (include/hesship.h HH_J2C_*)
J2C_CD, J2C_EIG = 0, 1


def _host_occ(mo_occ):
    if isinstance(mo_occ, torch.Tensor):
        mo_occ = mo_occ.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(mo_occ, dtype=np.float64))


def pack_triu_tilde(dm):
    """density_matrices.rs:61-77: [nao, nao, ...] -> [nao_tp, ...], off-diagonals x2."""
    nao = dm.shape[0]
    tail = tuple(dm.shape[2:])
    nmat = int(np.prod(tail)) if tail else 1
    d = fortran(dm.reshape(nao, nao, nmat))
    out = empty_f((nao * (nao + 1) // 2, nmat), dm.device)
    check(lib().hh_pack_triu_tilde(d.data_ptr(), nao, nmat, out.data_ptr(), stream_of(dm.device)))
    return out.reshape((nao * (nao + 1) // 2,) + tail)


def unpack_triu(tp, nao):
    """rstsr unpack_tri(Upper, Sy): [nao_tp, ...] -> [nao, nao, ...]."""
    tail = tuple(tp.shape[1:])
    nmat = int(np.prod(tail)) if tail else 1
    t = fortran(tp.reshape(tp.shape[0], nmat))
    out = empty_f((nao, nao, nmat), tp.device)
    check(lib().hh_unpack_triu(t.data_ptr(), nao, nmat, out.data_ptr(), stream_of(tp.device)))
    return out.reshape((nao, nao) + tail)


class _RIJKBase:
    nspin = 1

    def __init__(self, cderi, scale_j, scale_k, nao=None, _j2c=None, _threshold=1e-14, _policy="cd"):
        nao_tp, naux = cderi.shape
        if nao is None:
            nao = int(round((np.sqrt(8 * nao_tp + 1) - 1) / 2))
        assert nao * (nao + 1) // 2 == nao_tp, "cderi first dim must be nao*(nao+1)/2"
        self.nao, self.naux = nao, naux
        self.scale_j, self.scale_k = float(scale_j), float(scale_k)
        self.device = cderi.device
        cd = fortran(cderi)
        h = C.c_void_p()
        if _j2c is None:
            check(lib().hh_rijk_create(cd.data_ptr(), nao, naux, self.scale_j, self.scale_k, self.nspin,
                                       stream_of(self.device), C.byref(h)))
        else:
            j2 = fortran(_j2c)  # held across the call
            pol = {"cd": J2C_CD, "eig": J2C_EIG}[_policy]
            thr = 0.0 if _threshold is None else float(_threshold)  # None: no threshold
            check(lib().hh_rijk_create_from_j3c_decomp(cd.data_ptr(), j2.data_ptr(), nao, naux, pol, thr,
                                                       self.scale_j, self.scale_k, self.nspin,
                                                       stream_of(self.device), C.byref(h)))
        self._h = h
        self.nocc = None

    @classmethod
    def from_j3c(cls, j3c_tp, j2c, scale_j, scale_k, threshold=1e-14, policy="cd"):
        """cderi generated on the device from j3c [nao_tp, naux] and j2c (generate_cderi_with_decomp,
        hess_r.rs:1725-1735).  policy "cd" (decomp_j2c_cd, pure_decompose.rs:28-55: Cholesky, and the
        eigenvalue floor at `threshold` when it fails; threshold None: raise instead) or "eig"
        (decomp_j2c_eig, pure_decompose.rs:11-20: J^-1/2 over eigenvalues >= threshold, default
        1e-13).  The defaults are RHessRIJK::new_without_cderi's option (hess_r.rs:1752)."""
        return cls(j3c_tp, scale_j, scale_k, _j2c=j2c, _threshold=threshold, _policy=policy)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().hh_rijk_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_batch(self, nprop_batch):
        """Perturbations per K-contraction batch (0 = sized from memory)."""
        check(lib().hh_rijk_set_batch(self._h, int(nprop_batch)))

    def set_k_path(self, path):
        """BF16x3 RI-K algorithm: "fused" (operand generated inside the contraction, default where
        supported) or "materialized" (T generation + NT contraction)."""
        check(lib().hh_rijk_set_k_path(self._h, {"fused": 0, "materialized": 2}[path]))

    def set_response_path(self, path, budget_bytes=0):
        """Response algorithm: "direct" (per call, the reference's algebra; default) or "resident"
        (RHF and UHF, FP64 / BF16x3: the response matrix built once per prepare_response, one
        product per call and spin; falls back to direct beyond budget_bytes)."""
        check(lib().hh_rijk_set_response_path(self._h, {"direct": 0, "resident": 1}[path], int(budget_bytes)))

    def set_profiling(self, on=True):
        check(lib().hh_rijk_set_profiling(self._h, int(bool(on))))

    def timing(self):
        """dict of accumulated stage times (ms) and hot-kernel launch count."""
        out = (C.c_double * 5)()
        check(lib().hh_rijk_get_timing(self._h, out))
        return dict(k_hot_ms=out[0], k_total_ms=out[1], j_total_ms=out[2], response_ms=out[3],
                    k_hot_launches=int(out[4]))

    def reset_timing(self):
        check(lib().hh_rijk_reset_timing(self._h))

    def _prepare(self, mo_coeff, mo_occ):
        cs = [fortran(c) for c in mo_coeff]
        occ = [_host_occ(o) for o in mo_occ]
        nmo = (C.c_int64 * self.nspin)(*[c.shape[1] for c in cs])
        check(lib().hh_rijk_prepare_response(self._h, ptr_array([c.data_ptr() for c in cs]),
                                             ptr_array([o.ctypes.data for o in occ]), nmo))
        self.nocc = [int((o > 0).sum()) for o in occ]
        self._keep = (cs, occ)

    def _bras(self, bra):
        out, shapes = [], []
        for s, b in enumerate(bra):
            shapes.append(tuple(b.shape))
            out.append(fortran(b.reshape(self.nao, self.nocc[s], -1)))
        nprop = out[0].shape[2]
        assert all(o.shape[2] == nprop for o in out), "bra trailing dim (nprop) must agree across spins"
        return out, shapes, nprop

    def _response(self, bra, precision):
        bras, shapes, nprop = self._bras(bra)
        outs = [empty_f((self.nao, self.nocc[s], nprop), self.device) for s in range(self.nspin)]
        check(lib().hh_rijk_get_response_bra(self._h, ptr_array([b.data_ptr() for b in bras]), nprop,
                                             ptr_array([o.data_ptr() for o in outs]), PREC[precision]))
        return [o.reshape(sh) for o, sh in zip(outs, shapes)]

    def response_bra_separated(self, bra, do_j=True, do_k=True, precision="fp64"):
        bras, shapes, nprop = self._bras(bra)
        j_ao = empty_f((self.nao, self.nao, nprop), self.device) if do_j else None
        ks = [empty_f((self.nao, self.nocc[s], nprop), self.device) for s in range(self.nspin)] if do_k else []
        check(lib().hh_rijk_response_bra_separated(
            self._h, ptr_array([b.data_ptr() for b in bras]), nprop,
            j_ao.data_ptr() if do_j else None,
            ptr_array([k.data_ptr() for k in ks]) if do_k else None, PREC[precision]))
        return j_ao, [k.reshape((self.nao,) + sh[1:]) for k, sh in zip(ks, shapes)]


INT3C_KINDS = {"ip1": 0, "ip2": 1, "ipip1": 2, "ipvip1": 3, "ip1ip2": 4, "ipip2": 5}
_INT3C_FN = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_int32, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p)


def _skeleton(h, device, nao, mo_coeff, mo_occ, int3c, j2c_ip1, j2c_ipip1, j2c_ip1ip2, aux_ranges, aoslices,
              auxslices, j2c=None, with_deriv1=True):
    """hh_rijk_skeleton over device tensors.  int3c(kind, p0, p1) returns the derivative integrals
    int3c2e_<kind> for auxiliary functions [p0, p1) as a device tensor [nao, nao, p1-p0, 3] or
    [nao, nao, p1-p0, 3, 3] (kind in INT3C_KINDS)."""
    names = {v: k for k, v in INT3C_KINDS.items()}
    keep = []
    err = []

    def cb(_ctx, kind, p0, p1, out, stream):
        try:
            t = fortran(int3c(names[kind], int(p0), int(p1)))
            ncomp = 3 if kind < 2 else 9
            if t.numel() != nao * nao * (p1 - p0) * ncomp:
                raise ValueError(f"int3c({names[kind]}) returned {tuple(t.shape)}")
            keep.append(t)
            check(lib().hh_util_copy(out, t.data_ptr(), t.numel() * 8, stream))
            return 0
        except Exception as e:  # reported after the call returns
            err.append(e)
            return 1

    fn = _INT3C_FN(cb)
    cs = [fortran(c) for c in mo_coeff]
    occ = [_host_occ(o) for o in mo_occ]
    nmo = (C.c_int64 * len(cs))(*[c.shape[1] for c in cs])
    nocc = [int((o > 0).sum()) for o in occ]
    natm = len(aoslices)
    rng = (C.c_int64 * (2 * len(aux_ranges)))(*[int(v) for p in aux_ranges for v in p])
    asl = (C.c_int64 * (2 * natm))(*[int(v) for p in aoslices for v in p])
    xsl = (C.c_int64 * (2 * natm))(*[int(v) for p in auxslices for v in p])
    v1, v2, v3 = fortran(j2c_ip1), fortran(j2c_ipip1), fortran(j2c_ip1ip2)
    j2 = fortran(j2c) if j2c is not None else None
    de = empty_f((3, 3, natm, natm), device)
    d1 = [empty_f((nao, no, 3, natm), device) for no in nocc] if with_deriv1 else None
    rc = lib().hh_rijk_skeleton(h, ptr_array([c.data_ptr() for c in cs]), nmo, ptr_array([o.ctypes.data for o in occ]),
                                j2.data_ptr() if j2 is not None else None, v1.data_ptr(), v2.data_ptr(), v3.data_ptr(),
                                C.cast(fn, C.c_void_p), None, len(aux_ranges), rng, natm, asl, xsl, de.data_ptr(),
                                ptr_array([t.data_ptr() for t in d1]) if d1 is not None else None)
    if err:
        raise err[0]
    check(rc)
    return de, d1


class RHessRIJK(_RIJKBase):
    """RHessRIJK (hess_r.rs:1737-1954), response side, built from an existing cderi
    (new_with_cderi, hess_r.rs:1775-1794)."""
    nspin = 1

    def make_response_preparation(self, mo_coeff, mo_occ):
        self._prepare([mo_coeff], [mo_occ])

    def get_response_bra(self, bra, precision="fp64"):
        return self._response([bra], precision)[0]

    def make_skeleton(self, mo_coeff, mo_occ, int3c, j2c_ip1, j2c_ipip1, j2c_ip1ip2, aux_ranges, aoslices,
                      auxslices, j2c=None, with_deriv1=True):
        """(make_skeleton_hess, get_deriv1_bra) of hess_r.rs:1858-1909: de [3, 3, natm, natm] and
        deriv1_bra [nao, nocc, 3, natm] (None if with_deriv1 is False)."""
        de, d1 = _skeleton(self._h, self.device, self.nao, [mo_coeff], [mo_occ], int3c, j2c_ip1, j2c_ipip1,
                           j2c_ip1ip2, aux_ranges, aoslices, auxslices, j2c, with_deriv1)
        return de, (d1[0] if d1 is not None else None)


class UHessRIJK(_RIJKBase):
    """UHessRIJK (hess_u.rs:38-316), response side."""
    nspin = 2

    def make_response_preparation(self, mo_coeff, mo_occ):
        self._prepare(list(mo_coeff), list(mo_occ))

    def get_response_bra(self, bra, precision="fp64"):
        return self._response(list(bra), precision)

    def make_skeleton(self, mo_coeff, mo_occ, int3c, j2c_ip1, j2c_ipip1, j2c_ip1ip2, aux_ranges, aoslices,
                      auxslices, j2c=None, with_deriv1=True):
        """(make_skeleton_hess, get_deriv1_bra) of hess_u.rs:150-261: de and [deriv1_bra_a, _b]."""
        return _skeleton(self._h, self.device, self.nao, list(mo_coeff), list(mo_occ), int3c, j2c_ip1, j2c_ipip1,
                         j2c_ip1ip2, aux_ranges, aoslices, auxslices, j2c, with_deriv1)


def get_rijk_response_bra_separated(cderi, mo_coeff, mo_occ, bra, do_j, do_k, nbatch_aux=72, precision="fp64"):
    """hess_r.rs:82-186.  Returns (j_ao [nao,nao,nprop] or None, [k_bra_s]).

    `nbatch_aux` is accepted for signature parity; the GPU path chooses its own aux chunking."""
    nset = len(mo_coeff)
    cls = RHessRIJK if nset == 1 else UHessRIJK
    obj = cls(cderi, 1.0, 1.0)
    obj._prepare(list(mo_coeff), list(mo_occ))
    j_ao, ks = obj.response_bra_separated(list(bra), do_j, do_k, precision)
    if j_ao is not None:
        j_ao = j_ao.reshape((obj.nao, obj.nao) + tuple(bra[0].shape[2:]))
    return j_ao, ks
