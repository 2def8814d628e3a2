"""Oracle: DFT skeleton-Hessian setup passes over a fixed grid (NumPy FP64).  Test
infrastructure only.

Restates (ajz34/rstsr-showcase-hessian):
  src/numint_matmul/hess_rks.rs:71-172   make_hessian_setup (ao_dm0, rho, drho, de_fxc,
                                         de_vxc_diag, de_vxc_off, vmat_ip, vmat_deriv1)
  src/numint_matmul/hess_rks.rs:174-218  get_rho_vxc_fxc (the density part; the functional
                                         values come from the caller, LibXC being absent)
  src/numint_matmul/hess_rks.rs:220-279  get_drho
  src/numint_matmul/hess_rks.rs:281-300  get_de_fxc
  src/numint_matmul/hess_rks.rs:302-352  get_de_vxc_diag
  src/numint_matmul/hess_rks.rs:354-419  get_de_vxc_off
  src/numint_matmul/hess_rks.rs:421-469  get_vmat_ip
  src/numint_matmul/hess_rks.rs:471-492  get_vmat_vxc
  src/numint_matmul/hess_rks.rs:494-565  get_vmat_fxc, get_vmat_deriv1
  src/numint_matmul/hess_rks.rs:611-750  make_hessian_setup_batched (sums over grid chunks)
  src/numint_matmul/hess_uks.rs:56-384   the unrestricted counterparts
  src/numint_matmul/hess_rks.rs:866-872, hess_uks.rs:705-720  make_skeleton_hess sums

AO components (hess_rks.rs:7-29): 0 value; 1-3 x, y, z; 4-9 xx, xy, xz, yy, yz, zz; 10-19
xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz (sorted index tuples in lexicographic order).
ao [ng, nao, ncomp]; dm0 [nao, nao]; vxc_eff [ng, nvar]; fxc_eff [ng, nvar, nvar] (UKS:
[ng, nvar, 2] and [ng, nvar, 2, nvar, 2]); aoslices: list of (p0, p1) AO ranges per atom.
"""
import itertools

import numpy as np

NVAR = {"rho": 1, "sigma": 4, "tau": 5}
NCOMP_AO_DM0 = {"rho": 1, "sigma": 4, "tau": 4}    # hess_rks.rs:47-54
AO_DERIV = {"rho": 2, "sigma": 3, "tau": 3}        # hess_rks.rs:38-45
NCOMP_AO = {2: 10, 3: 20}
PAIRS6 = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]  # xx xy xz yy yz zz


def comp(*axes):
    """AO component index of the derivative along `axes` (0, 1, 2 = x, y, z)."""
    n = len(axes)
    if n == 0:
        return 0
    tuples = list(itertools.combinations_with_replacement(range(3), n))
    return sum(len(list(itertools.combinations_with_replacement(range(3), m))) for m in range(n)) + \
        tuples.index(tuple(sorted(axes)))


def ao_dm0(ao, dm0, den):
    """hess_rks.rs:113-115: ao_dm0[g, mu, c] = sum_nu ao[g, nu, c] dm0[nu, mu], c < ncomp."""
    nc = NCOMP_AO_DM0[den]
    return np.einsum("gvc,vu->guc", ao[:, :, :nc], dm0)


def get_rho(den, ao, aod):
    """hess_rks.rs:187-200."""
    ng = ao.shape[0]
    rho = np.zeros((ng, NVAR[den]))
    rho[:, 0] = np.sum(ao[:, :, 0] * aod[:, :, 0], axis=1)
    if den in ("sigma", "tau"):
        for t in range(1, 4):
            rho[:, t] = 2.0 * np.sum(ao[:, :, t] * aod[:, :, 0], axis=1)
    if den == "tau":
        rho[:, 4] = 0.5 * sum(np.sum(ao[:, :, t] * aod[:, :, t], axis=1) for t in range(1, 4))
    return rho


def drho_components(den):
    """(rho variable, direction t, bra component, ket component) of hess_rks.rs:235-263."""
    comps = [(0, t, comp(t), 0) for t in range(3)]
    if den in ("sigma", "tau"):
        comps += [(1 + r, t, comp(t, r), 0) for t in range(3) for r in range(3)]
        comps += [(1 + r, t, comp(t), comp(r)) for t in range(3) for r in range(3)]
    if den == "tau":
        comps += [(4, t, comp(t, r), comp(r)) for t in range(3) for r in range(3)]
    return comps


def get_drho(den, ao, aod, aoslices):
    """hess_rks.rs:220-279: drho [ng, nvar, 3, natm] = d rho / d R_A (AO centres move)."""
    ng = ao.shape[0]
    nvar = NVAR[den]
    drho = np.zeros((ng, nvar, 3, len(aoslices)))
    for A, (p0, p1) in enumerate(aoslices):
        for v, t, cb, ck in drho_components(den):
            drho[:, v, t, A] -= np.sum(ao[:, p0:p1, cb] * aod[:, p0:p1, ck], axis=1)
    drho[:, :1 if den == "rho" else 4] *= 2.0
    return drho


def get_de_fxc(wf, drho):
    """hess_rks.rs:281-300: de[t, s, A, B] = sum_g sum_xy wf[g,x,y] drho[g,x,t,A] drho[g,y,s,B]."""
    return np.einsum("gxy,gxta,gysb->tsab", wf, drho, drho)


def make_dao_vxc_diag(den, ao, aod, wv):
    """hess_rks.rs:302-339 (hess_rks_becke.rs:321-369): the per-AO kernel [nao, 6] over PAIRS6."""
    nao = ao.shape[1]
    dao = np.zeros((nao, 6))
    aow = aod[:, :, 0] * wv[:, 0:1]
    if den in ("sigma", "tau"):
        for r in range(1, 4):
            aow = aow + aod[:, :, r] * wv[:, r:r + 1]
    for k, (t, s) in enumerate(PAIRS6):
        dao[:, k] += 2.0 * np.sum(ao[:, :, comp(t, s)] * aow, axis=0)
    if den in ("sigma", "tau"):
        for k, (t, s) in enumerate(PAIRS6):
            a3 = sum(ao[:, :, comp(t, s, r)] * wv[:, 1 + r:2 + r] for r in range(3))
            dao[:, k] += 2.0 * np.sum(a3 * aod[:, :, 0], axis=0)
    if den == "tau":
        for r in range(3):
            aw = aod[:, :, r + 1] * wv[:, 4:5]
            for k, (t, s) in enumerate(PAIRS6):
                dao[:, k] += np.sum(ao[:, :, comp(r, t, s)] * aw, axis=0)
    return dao


def get_de_vxc_diag(den, ao, aod, wv, aoslices):
    """hess_rks.rs:302-352."""
    natm = len(aoslices)
    dao = make_dao_vxc_diag(den, ao, aod, wv)
    de6 = np.zeros((6, natm, natm))
    for A, (p0, p1) in enumerate(aoslices):
        de6[:, A, A] = dao[p0:p1].sum(axis=0)
    full = np.zeros((3, 3, natm, natm))
    for k, (t, s) in enumerate(PAIRS6):
        full[t, s] = full[s, t] = de6[k]
    return full


def make_dao_vxc_off(den, ao, wv):
    """hess_rks.rs:359-400 (before the transposition): [nao, nao, 3, 3]."""
    nao = ao.shape[1]
    dao = np.zeros((nao, nao, 3, 3))
    if den == "rho":
        for t in range(3):
            aowv = wv[:, 0:1] * ao[:, :, 1 + t]
            for s in range(3):
                dao[:, :, t, s] += aowv.T @ ao[:, :, 1 + s]
    else:
        for t in range(3):
            aowv = 0.5 * wv[:, 0:1] * ao[:, :, 1 + t]
            for r in range(3):
                aowv = aowv + wv[:, 1 + r:2 + r] * ao[:, :, comp(t, r)]
            for s in range(3):
                dao[:, :, t, s] += 2.0 * aowv.T @ ao[:, :, 1 + s]
    if den == "tau":
        tau = np.zeros((nao, nao, 3, 3))
        for r in range(3):
            for t in range(3):
                aowv = 0.5 * wv[:, 4:5] * ao[:, :, comp(t, r)]
                for s in range(t + 1):
                    tau[:, :, t, s] += aowv.T @ ao[:, :, comp(s, r)]
        for t in range(3):
            for s in range(t + 1):
                dao[:, :, t, s] += tau[:, :, t, s]
            for s in range(t):
                dao[:, :, s, t] += tau[:, :, t, s].T
    return dao


def get_de_vxc_off(den, ao, dm0, wv, aoslices):
    """hess_rks.rs:354-419."""
    dao = make_dao_vxc_off(den, ao, wv)
    dao = dao + dao.transpose(1, 0, 3, 2)
    natm = len(aoslices)
    de = np.zeros((3, 3, natm, natm))
    for A, (a0, a1) in enumerate(aoslices):
        for B, (b0, b1) in enumerate(aoslices):
            c = np.einsum("uvts,uv->ts", dao[a0:a1, b0:b1], dm0[a0:a1, b0:b1])
            de[:, :, A, B] = c
            de[:, :, B, A] = c.T
    return de


def get_vmat_ip(den, ao, wv):
    """hess_rks.rs:421-469: [nao, nao, 3]."""
    nao = ao.shape[1]
    v = np.zeros((nao, nao, 3))
    if den == "rho":
        aow = wv[:, 0:1] * ao[:, :, 0]
        for t in range(3):
            v[:, :, t] += ao[:, :, 1 + t].T @ aow
        return v
    aow = 0.5 * wv[:, 0:1] * ao[:, :, 0]
    for r in range(3):
        aow = aow + wv[:, 1 + r:2 + r] * ao[:, :, 1 + r]
    for t in range(3):
        v[:, :, t] += ao[:, :, 1 + t].T @ aow
    for t in range(3):
        aow_d = 0.5 * wv[:, 0:1] * ao[:, :, 1 + t]
        for r in range(3):
            aow_d = aow_d + wv[:, 1 + r:2 + r] * ao[:, :, comp(t, r)]
        v[:, :, t] += aow_d.T @ ao[:, :, 0]
    if den == "tau":
        for r in range(3):
            aow = 0.5 * wv[:, 4:5] * ao[:, :, 1 + r]
            for t in range(3):
                v[:, :, t] += ao[:, :, comp(t, r)].T @ aow
    return v


def get_vmat_vxc(vmat_ip, aoslices):
    """hess_rks.rs:471-492: [nao, nao, 3, natm]."""
    nao = vmat_ip.shape[0]
    out = np.zeros((nao, nao, 3, len(aoslices)))
    for A, (p0, p1) in enumerate(aoslices):
        out[p0:p1, :, :, A] -= vmat_ip[p0:p1]
    return out + out.transpose(1, 0, 2, 3)


def _vmat_fxc_from_wf_rho(den, ao, wf_rho_A, out_tA):
    """The per-atom body of hess_rks.rs:509-536: wf_rho_A [ng, nvar, 3] already scaled."""
    for t in range(3):
        if den == "rho":
            aow = wf_rho_A[:, 0, t:t + 1] * ao[:, :, 0]
        else:
            aow = sum(wf_rho_A[:, c, t:t + 1] * ao[:, :, c] for c in range(4))
        out_tA[:, :, t] += aow.T @ ao[:, :, 0]
        if den == "tau":
            for r in range(1, 4):
                out_tA[:, :, t] += (wf_rho_A[:, 4, t:t + 1] * ao[:, :, r]).T @ ao[:, :, r]


def _scale_wf_rho(den, wf_rho):
    wf_rho = wf_rho.copy()
    wf_rho[:, 0] *= 0.5
    if den == "tau":
        wf_rho[:, 4] *= 0.25
    return wf_rho


def get_vmat_fxc(den, ao, drho, wf, aoslices):
    """hess_rks.rs:494-543."""
    nao = ao.shape[1]
    natm = len(aoslices)
    out = np.zeros((nao, nao, 3, natm))
    for A in range(natm):
        wf_rho = _scale_wf_rho(den, np.einsum("gxy,gxt->gyt", wf, drho[:, :, :, A]))
        _vmat_fxc_from_wf_rho(den, ao, wf_rho, out[:, :, :, A])
    return out + out.transpose(1, 0, 2, 3)


def make_hessian_setup(den, ao, weights, dm0, aoslices, vxc, fxc):
    """hess_rks.rs:71-172 with the functional values given: the intermediates of one chunk."""
    aod = ao_dm0(ao, dm0, den)
    wv = weights[:, None] * vxc
    wf = weights[:, None, None] * fxc
    drho = get_drho(den, ao, aod, aoslices)
    vmat_ip = get_vmat_ip(den, ao, wv)
    return dict(
        de_fxc=get_de_fxc(wf, drho),
        de_vxc_diag=get_de_vxc_diag(den, ao, aod, wv, aoslices),
        de_vxc_off=get_de_vxc_off(den, ao, dm0, wv, aoslices),
        vmat_ip=vmat_ip,
        vmat_deriv1=get_vmat_fxc(den, ao, drho, wf, aoslices) + get_vmat_vxc(vmat_ip, aoslices),
    )


def make_hessian_setup_batched(den, ao, weights, dm0, aoslices, vxc, fxc, nchunk=None):
    """hess_rks.rs:611-750: the additive intermediates summed over grid chunks."""
    ng = ao.shape[0]
    nchunk = nchunk or ng
    out = None
    for g0 in range(0, ng, nchunk):
        s = slice(g0, min(ng, g0 + nchunk))
        r = make_hessian_setup(den, ao[s], weights[s], dm0, aoslices, vxc[s], fxc[s])
        out = r if out is None else {k: out[k] + r[k] for k in out}
    return out


def skeleton_hess(r):
    """hess_rks.rs:866-872."""
    return r["de_fxc"] + r["de_vxc_diag"] + r["de_vxc_off"]


def setup_rho(den, ao, dm0):
    """The density of get_rho_vxc_fxc (hess_rks.rs:187-200) for the caller's functional call."""
    return get_rho(den, ao, ao_dm0(ao, dm0, den))


# ---------------------------------------------------------------------------------------------
# Unrestricted (hess_uks.rs:56-384).  dm0 = (dm0_a, dm0_b); vxc [ng, nvar, 2];
# fxc [ng, nvar, 2, nvar, 2].

def setup_rho_uks(den, ao, dm0):
    """hess_uks.rs:78-97: rho [ng, nvar, 2]."""
    return np.stack([setup_rho(den, ao, d) for d in dm0], axis=2)


def get_de_fxc_uks(wf, drho):
    """hess_uks.rs:122-149: sum over the four spin blocks."""
    return sum(np.einsum("gxy,gxta,gysb->tsab", wf[:, :, s1, :, s2], drho[s1], drho[s2])
               for s1 in range(2) for s2 in range(2))


def get_vmat_fxc_uks(den, ao, drho, wf, aoslices):
    """hess_uks.rs:151-248: per output spin s2, wf_rho = sum_s1 wf[., s1, ., s2] drho[s1]."""
    nao = ao.shape[1]
    natm = len(aoslices)
    out = [np.zeros((nao, nao, 3, natm)) for _ in range(2)]
    for A in range(natm):
        for s2 in range(2):
            wf_rho = sum(np.einsum("gxy,gxt->gyt", wf[:, :, s1, :, s2], drho[s1][:, :, :, A]) for s1 in range(2))
            _vmat_fxc_from_wf_rho(den, ao, _scale_wf_rho(den, wf_rho), out[s2][:, :, :, A])
    return [o + o.transpose(1, 0, 2, 3) for o in out]


def make_hessian_setup_uks(den, ao, weights, dm0, aoslices, vxc, fxc):
    """hess_uks.rs:286-384 with the functional values given."""
    aod = [ao_dm0(ao, d, den) for d in dm0]
    wv = [weights[:, None] * vxc[:, :, s] for s in range(2)]
    wf = weights[:, None, None, None, None] * fxc
    drho = [get_drho(den, ao, a, aoslices) for a in aod]
    vmat_ip = [get_vmat_ip(den, ao, w) for w in wv]
    vf = get_vmat_fxc_uks(den, ao, drho, wf, aoslices)
    out = dict(de_fxc=get_de_fxc_uks(wf, drho))
    for s, tag in enumerate("ab"):
        out["de_vxc_diag_" + tag] = get_de_vxc_diag(den, ao, aod[s], wv[s], aoslices)
        out["de_vxc_off_" + tag] = get_de_vxc_off(den, ao, dm0[s], wv[s], aoslices)
        out["vmat_ip_" + tag] = vmat_ip[s]
        out["vmat_deriv1_" + tag] = vf[s] + get_vmat_vxc(vmat_ip[s], aoslices)
    return out


def skeleton_hess_uks(r):
    """hess_uks.rs:705-720."""
    return r["de_fxc"] + r["de_vxc_diag_a"] + r["de_vxc_off_a"] + r["de_vxc_diag_b"] + r["de_vxc_off_b"]
