"""Oracle: DFT skeleton-Hessian setup with the Becke grid shift (NumPy FP64).  Test
infrastructure only.

Restates (ajz34/rstsr-showcase-hessian):
  src/numint_matmul/hess_rks_becke.rs:127-141   quad_split_by_atom
  src/numint_matmul/hess_rks_becke.rs:610-700   get_de_becke_atom_1/2/3
  src/numint_matmul/hess_rks_becke.rs:702-784   contract_pvxc, get_de_becke_vxc_parts
  src/numint_matmul/hess_rks_becke.rs:786-836   xc_fock_stack
  src/numint_matmul/hess_rks_becke.rs:838-908   get_vmat_becke_parts
  src/numint_matmul/hess_rks_becke.rs:916-1053  make_hessian_setup_chunk_becke
  src/numint_matmul/hess_rks_becke.rs:1061-1266 symmetrize_ts_ab, make_hessian_setup_becke
  src/numint_matmul/hess_uks_becke.rs:240-684   the unrestricted terms and chunk
  src/numint_matmul/hess_uks_becke.rs:726-930   make_hessian_setup_becke_uks
The grid-fixed keys come from oracle.dft_setup (hess_rks.rs / hess_uks.rs), the Becke weight
derivatives from oracle.becke.  exc is the per-particle energy density (the energy density is
exc * rho[0]); vxc / fxc are the transformed kernels as in oracle.dft_setup.
"""
import numpy as np

from oracle import becke as obk
from oracle import dft_setup as ods

IDX_PAIR_TS = [0, 1, 2, 1, 3, 4, 2, 4, 5]   # hess_rks_becke.rs:762


def quad_split_by_atom(split, nsplit):
    """hess_rks_becke.rs:127-141: (atom, start, end) pieces of at most nsplit grids."""
    out = []
    for A in range(len(split) - 1):
        s = split[A]
        while s < split[A + 1]:
            e = min(s + nsplit, split[A + 1])
            out.append((A, s, e))
            s = e
    return out


def de_becke_atom_1(w, prho, fxc, drho):
    """hess_rks_becke.rs:630-637: [t, s, A] = sum_{g,x} w fxc_drho[g,x,t,A] prho[g,x,s]."""
    fd = np.einsum("gxy,gytA->gxtA", fxc, drho) * w[:, None, None, None]
    return np.einsum("gxtA,gxs->tsA", fd, prho)


def de_becke_atom_2(dwg, vxc, prho):
    """hess_rks_becke.rs:656-661: [t, s, A] = sum_{g,x} vxc[g,x] dw[g,t,A] prho[g,x,s]."""
    return np.einsum("gx,gtA,gxs->tsA", vxc, dwg, prho)


def de_becke_atom_3(w, prho, fxc):
    """hess_rks_becke.rs:679-687: [t, s] = sum_{g,x} fp[g,x,t] w prho[g,x,s]."""
    fp = np.einsum("gxy,gyt->gxt", fxc, prho)
    return np.einsum("gxt,gxs->ts", fp, prho * w[:, None, None])


def contract_pvxc(pvxc, atm_idx, aoslices):
    """hess_rks_becke.rs:702-712: row [t, s, B] of the chunk atom's column."""
    natm = len(aoslices)
    row = np.zeros((3, 3, natm))
    row[:, :, atm_idx] += 0.5 * pvxc.sum(axis=0)
    for B, (p0, p1) in enumerate(aoslices):
        row[:, :, B] -= pvxc[p0:p1].sum(axis=0)
    return row


def de_becke_vxc_parts(dao_diag, dao_off_sym, dm0, atm_idx, aoslices):
    """hess_rks_becke.rs:736-767 (dm0 / kernels may be per-spin lists: hess_uks_becke.rs:357-394)."""
    if isinstance(dao_diag, (list, tuple)):
        pd = sum(dao_diag)
        po = sum(np.einsum("uvts,uv->uts", o, d) for o, d in zip(dao_off_sym, dm0))
    else:
        pd = dao_diag
        po = np.einsum("uvts,uv->uts", dao_off_sym, dm0)
    nao = pd.shape[0]
    pd = pd[:, IDX_PAIR_TS].reshape(nao, 3, 3)
    return contract_pvxc(pd, atm_idx, aoslices), contract_pvxc(po, atm_idx, aoslices)


def xc_fock_stack(den, ao, wv):
    """hess_rks_becke.rs:786-836: fock [nao, nao, k] from grid-weighted fields wv [g, nvar, k]."""
    wv = wv.copy()
    wv[:, 0] *= 0.5
    nc = 1 if den == "rho" else 4
    aow = np.einsum("guc,gck->guk", ao[:, :, :nc], wv[:, :nc])
    x = np.einsum("gv,guk->vuk", ao[:, :, 0], aow)
    fock = x + x.transpose(1, 0, 2)
    if den == "tau":
        wv[:, 4] *= 0.5
        for j in range(1, 4):
            xj = np.einsum("gv,gu,gk->vuk", ao[:, :, j], ao[:, :, j], wv[:, 4])
            fock += xj.transpose(1, 0, 2)
    return fock


def _symm(x):
    """hess_rks_becke.rs:1061-1064: x + x.transpose(1, 0, 3, 2)."""
    return x + x.transpose(1, 0, 3, 2)


def _dao_off_sym(den, ao, wv):
    d = ods.make_dao_vxc_off(den, ao, wv)
    return d + d.transpose(1, 0, 3, 2)


def make_hessian_setup_chunk_becke(den, ao, weights, wquad, coords, atm_coords, adjustment, hardness, dm0,
                                   aoslices, atm_idx, exc, vxc, fxc):
    """hess_rks_becke.rs:916-1053 for one chunk of atom atm_idx's grids."""
    natm = len(aoslices)
    out = ods.make_hessian_setup(den, ao, weights, dm0, aoslices, vxc, fxc)
    aod = ods.ao_dm0(ao, dm0, den)
    rho0 = ods.get_rho(den, ao, aod)[:, 0]
    wv = weights[:, None] * vxc
    drho = ods.get_drho(den, ao, aod, aoslices)
    prho = -drho.sum(axis=3)
    ng = ao.shape[0]
    bk = obk.partition(coords, atm_coords, np.full(ng, atm_idx), wquad, adjustment, hardness, deriv=2,
                       contract_ddw=(exc * rho0)[None, :], output=(False, True, False))
    dwg = bk["dw"].transpose(2, 1, 0)                              # [g, t, A]
    vxc_drho = np.einsum("gxsB,gx->gsB", drho, vxc)
    out["de_becke_full_1"] = np.einsum("gtA,gsB->tsAB", dwg, vxc_drho)
    out["de_becke_full_2"] = bk["ddc"][..., 0].transpose(1, 3, 0, 2)
    out["de_becke_atom_1"] = de_becke_atom_1(weights, prho, fxc, drho)
    out["de_becke_atom_2"] = de_becke_atom_2(dwg, vxc, prho)
    out["de_becke_atom_3"] = de_becke_atom_3(weights, prho, fxc)
    dao_diag = ods.make_dao_vxc_diag(den, ao, aod, wv)
    vd, vo = de_becke_vxc_parts(dao_diag, _dao_off_sym(den, ao, wv), dm0, atm_idx, aoslices)
    out["de_becke_vxc_diag"], out["de_becke_vxc_off"] = vd, vo
    nao = ao.shape[1]
    vdw = np.zeros((nao, nao, 3, natm))
    for A in range(natm):
        vdw[:, :, :, A] = xc_fock_stack(den, ao, vxc[:, :, None] * dwg[:, None, :, A])
    out["vmat_becke_dw"] = vdw
    vip = out["vmat_ip"]
    out["vmat_becke_vxc"] = vip + vip.transpose(1, 0, 2)
    fxc_prho = np.einsum("gxy,gyt->gxt", fxc, prho)
    out["vmat_becke_fxc"] = xc_fock_stack(den, ao, fxc_prho * weights[:, None, None])
    return out


def make_hessian_setup_becke(den, ao, weights, wquad, coords, atm_coords, adjustment, hardness, dm0, aoslices,
                             split, exc, vxc, fxc, nchunk=None):
    """hess_rks_becke.rs:1106-1266: chunked accumulation, scatter, symmetrisation, assemblies."""
    natm = len(aoslices)
    nao = ao.shape[1]
    ng = ao.shape[0]
    pieces = quad_split_by_atom(split, nchunk or ng)
    acc = None
    col = dict(de_becke_atom_1=np.zeros((3, 3, natm, natm)), de_becke_atom_2=np.zeros((3, 3, natm, natm)),
               de_becke_atom_3=np.zeros((3, 3, natm, natm)), de_becke_vxc_diag=np.zeros((3, 3, natm, natm)),
               de_becke_vxc_off=np.zeros((3, 3, natm, natm)), vmat_becke_vxc=np.zeros((nao, nao, 3, natm)),
               vmat_becke_fxc=np.zeros((nao, nao, 3, natm)))
    for A, s0, s1 in pieces:
        s = slice(s0, s1)
        r = make_hessian_setup_chunk_becke(den, ao[s], weights[s], wquad[s], coords[s], atm_coords, adjustment,
                                           hardness, dm0, aoslices, A, exc[s], vxc[s], fxc[s])
        for k in list(col):
            if k == "de_becke_atom_3":
                col[k][:, :, A, A] += r.pop(k)
            else:
                col[k][..., A] += r.pop(k)
        acc = r if acc is None else {k: acc[k] + r[k] for k in acc}
    acc.update(col)
    for k in ("de_becke_full_1", "de_becke_atom_1", "de_becke_atom_2", "de_becke_vxc_diag", "de_becke_vxc_off"):
        acc[k] = _symm(acc[k])
    acc["de_xc_skeleton_no_becke"] = acc["de_vxc_diag"] + acc["de_vxc_off"] + acc["de_fxc"]
    acc["de_xc_skeleton"] = acc["de_xc_skeleton_no_becke"] + sum(
        acc[k] for k in ("de_becke_full_1", "de_becke_full_2", "de_becke_atom_1", "de_becke_atom_2",
                         "de_becke_atom_3", "de_becke_vxc_diag", "de_becke_vxc_off"))
    acc["vmat_deriv1_grid"] = (acc["vmat_deriv1"] + acc["vmat_becke_dw"] + acc["vmat_becke_vxc"]
                               + acc["vmat_becke_fxc"])
    return acc


# ---------------------------------------------------------------------------------------------
# Unrestricted (hess_uks_becke.rs).  dm0 = (dm0_a, dm0_b); vxc [ng, nvar, 2];
# fxc [ng, nvar, 2, nvar, 2].

def make_hessian_setup_chunk_becke_uks(den, ao, weights, wquad, coords, atm_coords, adjustment, hardness, dm0,
                                       aoslices, atm_idx, exc, vxc, fxc):
    """hess_uks_becke.rs:501-682 for one chunk of atom atm_idx's grids."""
    natm = len(aoslices)
    out = ods.make_hessian_setup_uks(den, ao, weights, dm0, aoslices, vxc, fxc)
    aod = [ods.ao_dm0(ao, d, den) for d in dm0]
    rho0 = sum(ods.get_rho(den, ao, a)[:, 0] for a in aod)
    wv = [weights[:, None] * vxc[:, :, s] for s in range(2)]
    drho = [ods.get_drho(den, ao, a, aoslices) for a in aod]
    prho = [-d.sum(axis=3) for d in drho]
    ng = ao.shape[0]
    bk = obk.partition(coords, atm_coords, np.full(ng, atm_idx), wquad, adjustment, hardness, deriv=2,
                       contract_ddw=(exc * rho0)[None, :], output=(False, True, False))
    dwg = bk["dw"].transpose(2, 1, 0)
    vxc_drho = sum(np.einsum("gxsB,gx->gsB", drho[s], vxc[:, :, s]) for s in range(2))
    out["de_becke_full_1"] = np.einsum("gtA,gsB->tsAB", dwg, vxc_drho)
    out["de_becke_full_2"] = bk["ddc"][..., 0].transpose(1, 3, 0, 2)
    out["de_becke_atom_1"] = sum(de_becke_atom_1(weights, prho[s1], fxc[:, :, s1, :, s2], drho[s2])
                                 for s1 in range(2) for s2 in range(2))
    out["de_becke_atom_2"] = sum(de_becke_atom_2(dwg, vxc[:, :, s], prho[s]) for s in range(2))
    a3 = np.zeros((3, 3))
    for s1 in range(2):
        for s2 in range(2):
            fp = np.einsum("gxy,gyt->gxt", fxc[:, :, s1, :, s2], prho[s2])
            a3 += np.einsum("gxt,gxs->ts", fp, prho[s1] * weights[:, None, None])
    out["de_becke_atom_3"] = a3
    dd = [ods.make_dao_vxc_diag(den, ao, aod[s], wv[s]) for s in range(2)]
    do = [_dao_off_sym(den, ao, wv[s]) for s in range(2)]
    out["de_becke_vxc_diag"], out["de_becke_vxc_off"] = de_becke_vxc_parts(dd, do, dm0, atm_idx, aoslices)
    nao = ao.shape[1]
    for s, tag in enumerate("ab"):
        vdw = np.zeros((nao, nao, 3, natm))
        for A in range(natm):
            vdw[:, :, :, A] = xc_fock_stack(den, ao, vxc[:, :, s, None] * dwg[:, None, :, A])
        out["vmat_becke_dw_" + tag] = vdw
        vip = out["vmat_ip_" + tag]
        out["vmat_becke_vxc_" + tag] = vip + vip.transpose(1, 0, 2)
        fp = sum(np.einsum("gxy,gyt->gxt", fxc[:, :, s, :, s2], prho[s2]) for s2 in range(2))
        out["vmat_becke_fxc_" + tag] = xc_fock_stack(den, ao, fp * weights[:, None, None])
    return out


def make_hessian_setup_becke_uks(den, ao, weights, wquad, coords, atm_coords, adjustment, hardness, dm0, aoslices,
                                 split, exc, vxc, fxc, nchunk=None):
    """hess_uks_becke.rs:726-930."""
    natm = len(aoslices)
    nao = ao.shape[1]
    ng = ao.shape[0]
    pieces = quad_split_by_atom(split, nchunk or ng)
    keys4 = ("de_becke_atom_1", "de_becke_atom_2", "de_becke_atom_3", "de_becke_vxc_diag", "de_becke_vxc_off")
    col = {k: np.zeros((3, 3, natm, natm)) for k in keys4}
    for k in ("vmat_becke_vxc_a", "vmat_becke_vxc_b", "vmat_becke_fxc_a", "vmat_becke_fxc_b"):
        col[k] = np.zeros((nao, nao, 3, natm))
    acc = None
    for A, s0, s1 in pieces:
        s = slice(s0, s1)
        r = make_hessian_setup_chunk_becke_uks(den, ao[s], weights[s], wquad[s], coords[s], atm_coords, adjustment,
                                               hardness, dm0, aoslices, A, exc[s], vxc[s], fxc[s])
        for k in list(col):
            if k == "de_becke_atom_3":
                col[k][:, :, A, A] += r.pop(k)
            else:
                col[k][..., A] += r.pop(k)
        acc = r if acc is None else {k: acc[k] + r[k] for k in acc}
    acc.update(col)
    for k in ("de_becke_full_1", "de_becke_atom_1", "de_becke_atom_2", "de_becke_vxc_diag", "de_becke_vxc_off"):
        acc[k] = _symm(acc[k])
    acc["de_xc_skeleton_no_becke"] = ods.skeleton_hess_uks(acc)
    acc["de_xc_skeleton"] = acc["de_xc_skeleton_no_becke"] + sum(
        acc[k] for k in ("de_becke_full_1", "de_becke_full_2", "de_becke_atom_1", "de_becke_atom_2",
                         "de_becke_atom_3", "de_becke_vxc_diag", "de_becke_vxc_off"))
    for tag in "ab":
        acc["vmat_deriv1_grid_" + tag] = (acc["vmat_deriv1_" + tag] + acc["vmat_becke_dw_" + tag]
                                          + acc["vmat_becke_vxc_" + tag] + acc["vmat_becke_fxc_" + tag])
    return acc
