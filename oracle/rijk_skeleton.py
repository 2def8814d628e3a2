"""Oracle: RI-JK skeleton Hessian and first-derivative Fock terms (NumPy FP64).  Test
infrastructure only.

Restates the naive forms of (ajz34/rstsr-showcase-hessian):
  src/ri_jk/hess_r_naive.rs:46-315   get_decomposed_rij_skeleton_deriv2_naive  (de_J20_*, de_J11_*, de_J02_*)
  src/ri_jk/hess_r_naive.rs:317-652  get_decomposed_rik_skeleton_deriv2_naive  (de_K20_*, de_K11_*, de_K02_*)
  src/ri_jk/hess_r_naive.rs:654-752  get_rij_deriv1_ao_naive                   (j1ao_aux0, j1ao_aux1)
  src/ri_jk/hess_r_naive.rs:754-860  get_rik_deriv1_ao_naive                   (k1ao_aux0, k1ao_aux1)
and the consumers that scale and sum them:
  src/ri_jk/hess_r.rs:1858-1909      RHessRIJK::make_skeleton_hess / get_deriv1_bra (RHF)
  src/ri_jk/hess_u.rs:150-261        UHessRIJK::make_skeleton_hess / get_deriv1_bra (UHF)
The optimized path (hess_r.rs:202-1718) computes the same per-key quantities.

Integrals (dict `I`, libcint layouts as the reference reads them, hess_intor reshaping a
9-component derivative to [.., 3, 3] column-major, cint_handling.rs:20-44):
  j3c [nao, nao, naux]; ip1, ip2 [nao, nao, naux, 3]; ipip1, ipvip1, ip1ip2, ipip2
  [nao, nao, naux, 3, 3]; j2c [naux, naux]; j2c_ip1 [naux, naux, 3]; j2c_ipip1, j2c_ip1ip2
  [naux, naux, 3, 3].  aoslices / auxslices: (p0, p1) per listed atom.
The einsum subscripts below are the reference's (written for column-major tensors; einsum is
layout-free).
"""
import numpy as np


def _es(sub, *ops):
    return np.einsum(sub, *ops, optimize=True)


def _pair_blocks(dbas, sl_a, sl_b, scale, de):
    """de[.., .., B, A] += scale * sum over the (A, B) block of dbas[u-like, v-like, s, t]."""
    for A, (a0, a1) in enumerate(sl_a):
        for B, (b0, b1) in enumerate(sl_b):
            de[:, :, B, A] += scale * dbas[a0:a1, b0:b1].sum(axis=(0, 1))


def _diag_blocks(dbas, sl, scale, de, naxes):
    for A, (a0, a1) in enumerate(sl):
        de[:, :, A, A] += scale * dbas[a0:a1].sum(axis=tuple(range(naxes)))


def _sym(de):
    return de + de.transpose(1, 0, 3, 2)


def rij_skeleton_naive(I, dm0, aoslices, auxslices):
    """hess_r_naive.rs:46-315."""
    natm = len(aoslices)
    z = lambda: np.zeros((3, 3, natm, natm))
    Vi = np.linalg.inv(I["j2c"])
    j3c, ip1, ip2 = I["j3c"], I["ip1"], I["ip2"]
    j2c_ip1 = I["j2c_ip1"]
    D = dm0
    out = {}
    # J20
    d = z(); _pair_blocks(_es("uvPt,PQ,klQs,uv,kl->ukst", ip1, Vi, ip1, D, D), aoslices, aoslices, 4.0, d)
    out["de_J20_1"] = d
    d = z(); _pair_blocks(_es("uvPst,PQ,klQ,uv,kl->uvst", I["ipvip1"], Vi, j3c, D, D), aoslices, aoslices, 2.0, d)
    out["de_J20_2"] = d
    d = z(); _diag_blocks(_es("uvPst,PQ,klQ,uv,kl->uvst", I["ipip1"], Vi, j3c, D, D), aoslices, 2.0, d, 2)
    out["de_J20_3"] = d
    # J11
    d = z(); _pair_blocks(_es("uvPst,PQ,klQ,uv,kl->uPst", I["ip1ip2"], Vi, j3c, D, D), aoslices, auxslices, 2.0, d)
    out["de_J11_1"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,QRs,RS,klS,uv,kl->uRst", ip1, Vi, j2c_ip1, Vi, j3c, D, D), aoslices,
                          auxslices, 2.0, d)
    out["de_J11_2"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,QRs,RS,klS,uv,kl->uQst", ip1, Vi, j2c_ip1, Vi, j3c, D, D), aoslices,
                          auxslices, -2.0, d)
    out["de_J11_3"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,klQs,uv,kl->uQst", ip1, Vi, ip2, D, D), aoslices, auxslices, 2.0, d)
    out["de_J11_4"] = _sym(d)
    # J02
    d = z(); _diag_blocks(_es("uvPst,PQ,klQ,uv,kl->Pst", I["ipip2"], Vi, j3c, D, D), auxslices, 1.0, d, 1)
    out["de_J02_1"] = d
    d = z(); _diag_blocks(_es("uvP,PQ,QRst,RS,klS,uv,kl->Qst", j3c, Vi, I["j2c_ipip1"], Vi, j3c, D, D), auxslices,
                          -1.0, d, 1)
    out["de_J02_2"] = d
    d = z(); _pair_blocks(_es("uvP,PQ,QRst,RS,klS,uv,kl->QRst", j3c, Vi, I["j2c_ip1ip2"], Vi, j3c, D, D), auxslices,
                          auxslices, -0.5, d)
    out["de_J02_3a"] = _sym(d)
    d = z(); _pair_blocks(_es("uvP,PQ,QRt,RS,STs,TU,klU,uv,kl->QTst", j3c, Vi, j2c_ip1, Vi, j2c_ip1, Vi, j3c, D, D),
                          auxslices, auxslices, -0.5, d)
    out["de_J02_3b"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,QRs,RS,klS,uv,kl->PQst", ip2, Vi, j2c_ip1, Vi, j3c, D, D), auxslices,
                          auxslices, -1.0, d)
    out["de_J02_4"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,klQs,uv,kl->PQst", ip2, Vi, ip2, D, D), auxslices, auxslices, 0.5, d)
    out["de_J02_5"] = _sym(d)
    d = z(); _pair_blocks(_es("uvP,PQ,RQt,RS,STs,TU,klU,uv,kl->RSst", j3c, Vi, j2c_ip1, Vi, j2c_ip1, Vi, j3c, D, D),
                          auxslices, auxslices, 0.5, d)
    out["de_J02_6"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,RQs,RS,klS,uv,kl->PRst", ip2, Vi, j2c_ip1, Vi, j3c, D, D), auxslices,
                          auxslices, -1.0, d)
    out["de_J02_7"] = _sym(d)
    d = z(); _pair_blocks(_es("uvP,PQ,QRt,RS,STs,TU,klU,uv,kl->RTst", j3c, Vi, j2c_ip1, Vi, j2c_ip1, Vi, j3c, D, D),
                          auxslices, auxslices, 1.0, d)
    out["de_J02_8"] = _sym(d)
    return out


def rik_skeleton_naive(I, mocc_2, aoslices, auxslices):
    """hess_r_naive.rs:317-652 (mocc_2 = mocc * sqrt(occ))."""
    natm = len(aoslices)
    z = lambda: np.zeros((3, 3, natm, natm))
    Vi = np.linalg.inv(I["j2c"])
    j3c, ip1, ip2 = I["j3c"], I["ip1"], I["ip2"]
    j2c_ip1 = I["j2c_ip1"]
    C = mocc_2
    out = {}
    # K20
    d = z(); _pair_blocks(_es("uvPt,PQ,klQs,ui,vj,ki,lj->ukst", ip1, Vi, ip1, C, C, C, C), aoslices, aoslices, 2.0, d)
    out["de_K20_1a"] = d
    d = z(); _pair_blocks(_es("uvPt,PQ,klQs,ui,vj,kj,li->ukst", ip1, Vi, ip1, C, C, C, C), aoslices, aoslices, 2.0, d)
    out["de_K20_1b"] = d
    d = z(); _pair_blocks(_es("uvPst,PQ,klQ,ui,vj,ki,lj->uvst", I["ipvip1"], Vi, j3c, C, C, C, C), aoslices, aoslices,
                          2.0, d)
    out["de_K20_2"] = d
    d = z(); _diag_blocks(_es("uvPst,PQ,klQ,ui,vj,ki,lj->uvst", I["ipip1"], Vi, j3c, C, C, C, C), aoslices, 2.0, d, 2)
    out["de_K20_3"] = d
    # K11
    d = z(); _pair_blocks(_es("uvPst,PQ,klQ,vi,li,uj,kj->uPst", I["ip1ip2"], Vi, j3c, C, C, C, C), aoslices,
                          auxslices, 2.0, d)
    out["de_K11_1"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,QRs,RS,klS,ui,vj,ki,lj->uRst", ip1, Vi, j2c_ip1, Vi, j3c, C, C, C, C),
                          aoslices, auxslices, 2.0, d)
    out["de_K11_2"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,QRs,RS,klS,ui,vj,ki,lj->uQst", ip1, Vi, j2c_ip1, Vi, j3c, C, C, C, C),
                          aoslices, auxslices, -2.0, d)
    out["de_K11_3"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,klQs,ui,vj,ki,lj->uQst", ip1, Vi, ip2, C, C, C, C), aoslices, auxslices, 2.0, d)
    out["de_K11_4"] = _sym(d)
    # K02
    d = z(); _diag_blocks(_es("uvPst,PQ,klQ,ui,vj,ki,lj->Pst", I["ipip2"], Vi, j3c, C, C, C, C), auxslices, 1.0, d, 1)
    out["de_K02_1"] = d
    d = z(); _diag_blocks(_es("uvP,PQ,QRst,RS,klS,ui,vj,ki,lj->Qst", j3c, Vi, I["j2c_ipip1"], Vi, j3c, C, C, C, C),
                          auxslices, -1.0, d, 1)
    out["de_K02_2"] = d
    d = z(); _pair_blocks(_es("uvP,PQ,QRst,RS,klS,ui,vj,ki,lj->QRst", j3c, Vi, I["j2c_ip1ip2"], Vi, j3c, C, C, C, C),
                          auxslices, auxslices, -0.5, d)
    out["de_K02_3a"] = _sym(d)
    d = z(); _pair_blocks(_es("uvP,PQ,QRt,RS,STs,TU,klU,ui,vj,ki,lj->QTst", j3c, Vi, j2c_ip1, Vi, j2c_ip1, Vi, j3c,
                              C, C, C, C), auxslices, auxslices, -0.5, d)
    out["de_K02_3b"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,QRs,RS,klS,ui,vj,ki,lj->PQst", ip2, Vi, j2c_ip1, Vi, j3c, C, C, C, C),
                          auxslices, auxslices, -1.0, d)
    out["de_K02_4"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,klQs,ui,vj,ki,lj->PQst", ip2, Vi, ip2, C, C, C, C), auxslices, auxslices, 0.5, d)
    out["de_K02_5"] = _sym(d)
    d = z(); _pair_blocks(_es("uvP,PQ,RQt,RS,STs,TU,klU,ui,vj,ki,lj->RSst", j3c, Vi, j2c_ip1, Vi, j2c_ip1, Vi, j3c,
                              C, C, C, C), auxslices, auxslices, 0.5, d)
    out["de_K02_6"] = _sym(d)
    d = z(); _pair_blocks(_es("uvPt,PQ,RQs,RS,klS,ui,vj,ki,lj->PRst", ip2, Vi, j2c_ip1, Vi, j3c, C, C, C, C),
                          auxslices, auxslices, -1.0, d)
    out["de_K02_7"] = _sym(d)
    d = z(); _pair_blocks(_es("uvP,PQ,QRt,RS,STs,TU,klU,ui,vj,ki,lj->QSst", j3c, Vi, j2c_ip1, Vi, j2c_ip1, Vi, j3c,
                              C, C, C, C), auxslices, auxslices, 1.0, d)
    out["de_K02_8"] = _sym(d)
    return out


def rij_deriv1_ao_naive(I, dm0, aoslices, auxslices):
    """hess_r_naive.rs:654-752: j1ao_aux0 + j1ao_aux1, [nao, nao, 3, natm]."""
    nao, natm = dm0.shape[0], len(aoslices)
    Vi = np.linalg.inv(I["j2c"])
    j3c, ip1, ip2, j2c_ip1 = I["j3c"], I["ip1"], I["ip2"], I["j2c_ip1"]
    D = dm0
    aux0 = np.zeros((nao, nao, 3, natm))
    scr1 = _es("uvPt,PQ,klQ,kl->uvt", ip1, Vi, j3c, D)
    for A, (p0, p1) in enumerate(aoslices):
        aux0[p0:p1, :, :, A] -= scr1[p0:p1]
        aux0[:, p0:p1, :, A] -= scr1[p0:p1].transpose(1, 0, 2)
        aux0[:, :, :, A] -= 2.0 * _es("klPt,PQ,uvQ,kl->uvt", ip1[p0:p1], Vi, j3c, D[p0:p1])
    aux1 = np.zeros((nao, nao, 3, natm))
    for A, (p0, p1) in enumerate(auxslices):
        s = slice(p0, p1)
        aux1[..., A] -= _es("uvPt,PQ,klQ,kl->uvt", ip2[:, :, s], Vi[s], j3c, D)
        aux1[..., A] -= _es("uvP,PQ,klQt,kl->uvt", j3c, Vi[:, s], ip2[:, :, s], D)
        aux1[..., A] += _es("uvP,PQ,QRt,RS,klS,kl->uvt", j3c, Vi[:, s], j2c_ip1[s], Vi, j3c, D)
        aux1[..., A] += _es("uvP,PQ,RQt,RS,klS,kl->uvt", j3c, Vi, j2c_ip1[s], Vi[s], j3c, D)
    return aux0 + aux1


def rik_deriv1_ao_naive(I, mocc_2, aoslices, auxslices):
    """hess_r_naive.rs:754-860: k1ao_aux0 + k1ao_aux1, [nao, nao, 3, natm]."""
    nao, natm = mocc_2.shape[0], len(aoslices)
    Vi = np.linalg.inv(I["j2c"])
    j3c, ip1, ip2, j2c_ip1 = I["j3c"], I["ip1"], I["ip2"], I["j2c_ip1"]
    C = mocc_2
    aux0 = np.zeros((nao, nao, 3, natm))
    scr1 = _es("uvPt,PQ,klQ,vi,li->ukt", ip1, Vi, j3c, C, C)
    for A, (p0, p1) in enumerate(aoslices):
        aux0[p0:p1, :, :, A] -= scr1[p0:p1]
        aux0[:, p0:p1, :, A] -= scr1[p0:p1].transpose(1, 0, 2)
        scr2 = _es("klPt,PQ,uvQ,ki,ui->lvt", ip1[p0:p1], Vi, j3c, C[p0:p1], C)
        aux0[:, :, :, A] -= scr2 + scr2.transpose(1, 0, 2)
    aux1 = np.zeros((nao, nao, 3, natm))
    for A, (p0, p1) in enumerate(auxslices):
        s = slice(p0, p1)
        aux1[..., A] -= _es("uvPt,PQ,klQ,vi,li->ukt", ip2[:, :, s], Vi[s], j3c, C, C)
        aux1[..., A] -= _es("uvP,PQ,klQt,vi,li->ukt", j3c, Vi[:, s], ip2[:, :, s], C, C)
        aux1[..., A] += _es("uvP,PQ,QRt,RS,klS,vi,li->ukt", j3c, Vi[:, s], j2c_ip1[s], Vi, j3c, C, C)
        aux1[..., A] += _es("uvP,PQ,RQt,RS,klS,vi,li->ukt", j3c, Vi, j2c_ip1[s], Vi[s], j3c, C, C)
    return aux0 + aux1


def _occupied(mo_coeff, mo_occ):
    idx = mo_occ > 0
    mocc = mo_coeff[:, idx]
    return mocc, mocc * np.sqrt(mo_occ[idx])[None, :]


def skeleton_rhf(I, mo_coeff, mo_occ, scale_j, scale_k, aoslices, auxslices):
    """hess_r.rs:1858-1909: (de_skeleton [3, 3, natm, natm], deriv1_bra [nao, nocc, 3, natm]) with
    de = scale_j de_J - 0.5 scale_k de_K and deriv1_bra = scale_j j1ao mocc - 0.5 scale_k k1ao mocc
    (k1bra is the right half-transform k1ao @ mocc)."""
    mocc, mocc_2 = _occupied(mo_coeff, mo_occ)
    dm0 = mocc_2 @ mocc_2.T
    natm = len(aoslices)
    de = np.zeros((3, 3, natm, natm))
    d1 = np.zeros((mocc.shape[0], mocc.shape[1], 3, natm))
    if scale_j != 0.0:
        de += scale_j * sum(rij_skeleton_naive(I, dm0, aoslices, auxslices).values())
        d1 += scale_j * _es("uvxa,vi->uixa", rij_deriv1_ao_naive(I, dm0, aoslices, auxslices), mocc)
    if scale_k != 0.0:
        de -= 0.5 * scale_k * sum(rik_skeleton_naive(I, mocc_2, aoslices, auxslices).values())
        d1 -= 0.5 * scale_k * _es("uvxa,vi->uixa", rik_deriv1_ao_naive(I, mocc_2, aoslices, auxslices), mocc)
    return de, d1


def skeleton_uhf(I, mo_coeff, mo_occ, scale_j, scale_k, aoslices, auxslices):
    """hess_u.rs:150-261: J from the total density, K per spin with coefficient -1;
    deriv1_bra[s] = scale_j j1ao mocc_s - scale_k k1ao_s mocc_s."""
    occ = [_occupied(c, o) for c, o in zip(mo_coeff, mo_occ)]
    dm0 = sum(m2 @ m2.T for _, m2 in occ)
    natm = len(aoslices)
    de = np.zeros((3, 3, natm, natm))
    d1 = [np.zeros((m.shape[0], m.shape[1], 3, natm)) for m, _ in occ]
    if scale_j != 0.0:
        de += scale_j * sum(rij_skeleton_naive(I, dm0, aoslices, auxslices).values())
        j1ao = rij_deriv1_ao_naive(I, dm0, aoslices, auxslices)
        for s, (m, _) in enumerate(occ):
            d1[s] += scale_j * _es("uvxa,vi->uixa", j1ao, m)
    if scale_k != 0.0:
        for s, (m, m2) in enumerate(occ):
            de -= scale_k * sum(rik_skeleton_naive(I, m2, aoslices, auxslices).values())
            d1[s] -= scale_k * _es("uvxa,vi->uixa", rik_deriv1_ao_naive(I, m2, aoslices, auxslices), m)
    return de, d1
