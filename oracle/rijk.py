"""Oracle: RI-JK response (FP64 NumPy restatement).  Test infrastructure only.

Reference files (ajz34/rstsr-showcase-hessian):
  src/ri_jk/hess_r.rs:82-186      get_rijk_response_bra_separated
  src/ri_jk/hess_r.rs:1916-1953   RHessRIJK::get_response_bra (RHF assembly)
  src/ri_jk/hess_u.rs:270-313     UHessRIJK::get_response_bra (UHF assembly)
  src/ri_jk/hess_r_naive.rs:861-903   naive RHF response (einsum form)
  src/ri_jk/hess_u_naive.rs:115-167   naive UHF response (einsum form)
  src/util/density_matrices.rs:48-77  pack_triu_tilde
  src/ri_jk/pure_decompose.rs:107-133 Cholesky of j2c (cderi = j3c L^-T)
"""
import numpy as np


def tri_index(nao):
    """Packed-upper (column-major) index of (mu, nu), mu <= nu: nu*(nu+1)/2 + mu."""
    mu, nu = np.triu_indices(nao)
    return mu, nu, nu * (nu + 1) // 2 + mu


def pack_triu(m):
    """Pack the upper triangle (mu <= nu) of m[nao, nao, ...] to [nao_tp, ...]."""
    nao = m.shape[0]
    mu, nu, tp = tri_index(nao)
    out = np.empty((nao * (nao + 1) // 2,) + m.shape[2:], dtype=m.dtype)
    out[tp] = m[mu, nu]
    return out


def pack_triu_tilde(dm):
    """density_matrices.rs:48-77: packed upper triangle, off-diagonals x2."""
    nao = dm.shape[0]
    mu, nu, tp = tri_index(nao)
    out = np.empty((nao * (nao + 1) // 2,) + dm.shape[2:], dtype=dm.dtype)
    w = np.where(mu == nu, 1.0, 2.0).reshape((-1,) + (1,) * (dm.ndim - 2))
    out[tp] = dm[mu, nu] * w
    return out


def unpack_triu(tp_arr, nao):
    """rstsr unpack_tri(Upper, Sy): symmetric matrix from the packed upper triangle."""
    mu, nu, tp = tri_index(nao)
    out = np.zeros((nao, nao) + tp_arr.shape[1:], dtype=tp_arr.dtype)
    out[mu, nu] = tp_arr[tp]
    out[nu, mu] = tp_arr[tp]
    return out


def mocc_of(mo_coeff, mo_occ):
    return mo_coeff[:, np.asarray(mo_occ) > 0]


def response_bra_separated(cderi, mo_coeff, mo_occ, bra, do_j, do_k, nbatch_aux=72):
    """hess_r.rs:82-186.  Returns (j_ao [nao,nao,nprop] or None, [k_bra_s [nao,nocc_s,nprop]]).

    Same loop structure as the reference: J from the summed, symmetrised,
    tilde-packed dm1 (two GEMMs over packed cderi); K per spin, per aux batch
    of `nbatch_aux`, per perturbation `a` (hess_r.rs:144-183).
    """
    nset = len(mo_coeff)
    nao = mo_coeff[0].shape[0]
    nao_tp, naux = cderi.shape
    assert nao_tp == nao * (nao + 1) // 2
    mocc = [mocc_of(mo_coeff[s], mo_occ[s]) for s in range(nset)]
    nocc = [m.shape[1] for m in mocc]
    bra = [np.asarray(bra[s]).reshape(nao, nocc[s], -1) for s in range(nset)]
    nprop = bra[0].shape[2]
    j_ao = None
    k_bras = []
    if do_j:
        dm1 = np.zeros((nao, nao, nprop))
        for s in range(nset):
            dm1 += np.einsum("uia,vi->uva", bra[s], mocc[s])
        dm1 = dm1 + dm1.transpose(1, 0, 2)
        dm1_tp = pack_triu_tilde(dm1)
        itm = cderi.T @ dm1_tp
        resp_tp = 2.0 * (cderi @ itm)
        j_ao = unpack_triu(resp_tp, nao)
    if do_k:
        for s in range(nset):
            mo, b = mocc[s], bra[s]
            no = nocc[s]
            resp = np.zeros_like(b)
            for p0 in range(0, naux, nbatch_aux):
                p1 = min(p0 + nbatch_aux, naux)
                resp += k_batch(cderi[:, p0:p1], mo, b)
            k_bras.append(resp)
    return j_ao, k_bras


def k_batch(cderi_batch, mo, b):
    """One aux batch of the K loop, hess_r.rs:148-182, as the reference's GEMM sequence.

    Returns the batch's contribution to k_bra [nao, nocc, nprop] (sign included)."""
    nao, no = mo.shape
    nb = cderi_batch.shape[1]
    nprop = b.shape[2]
    B = unpack_triu(cderi_batch, nao).reshape(nao, nao * nb)      # B[u, (v, P)]
    bxo = (B.T @ mo).reshape(nao, nb, no)                          # cderi_bxo[v, P, i]
    oxo = (mo.T @ bxo.reshape(nao, nb * no)).reshape(no, nb, no)   # cderi_oxo[j, P, i]
    out = np.zeros((nao, no, nprop))
    for a in range(nprop):
        bxo1 = (B.T @ b[:, :, a]).reshape(nao, nb, no)             # uPj, part 0
        out[:, :, a] -= bxo1.reshape(nao, nb * no) @ oxo.reshape(no, nb * no).T
        oxo1 = (mo.T @ bxo1.reshape(nao, nb * no)).reshape(no, nb, no)
        out[:, :, a] -= bxo.reshape(nao, nb * no) @ oxo1.reshape(no, nb * no).T
    return out


def rhf_response_bra(cderi, mo_coeff, mo_occ, bra, scale_j, scale_k, nbatch_aux=72):
    """hess_r.rs:1916-1953: resp = scale_j * J_ao @ mocc + scale_k * K_bra."""
    j_ao, k = response_bra_separated(cderi, [mo_coeff], [mo_occ], [bra],
                                     scale_j != 0.0, scale_k != 0.0, nbatch_aux)
    mocc = mocc_of(mo_coeff, mo_occ)
    nao, nocc = mocc.shape
    shape = np.asarray(bra).shape
    resp = np.zeros((nao, nocc, int(np.prod(shape[2:]))))
    if j_ao is not None:
        resp += scale_j * np.einsum("uva,vi->uia", j_ao, mocc)
    if k:
        resp += scale_k * k[0].reshape(nao, nocc, -1)
    return resp.reshape(shape)


def uhf_response_bra(cderi, mo_coeff, mo_occ, bra, scale_j, scale_k, nbatch_aux=72):
    """hess_u.rs:270-313: r_s = 0.5 * scale_j * J_ao @ mocc_s + scale_k * K_bra_s."""
    j_ao, k = response_bra_separated(cderi, mo_coeff, mo_occ, bra,
                                     scale_j != 0.0, scale_k != 0.0, nbatch_aux)
    out = []
    for s in range(2):
        mocc = mocc_of(mo_coeff[s], mo_occ[s])
        nao, nocc = mocc.shape
        shape = np.asarray(bra[s]).shape
        r = np.zeros((nao, nocc, int(np.prod(shape[2:]))))
        if j_ao is not None:
            r += 0.5 * scale_j * np.einsum("uva,vi->uia", j_ao, mocc)
        if k:
            r += scale_k * k[s].reshape(nao, nocc, -1)
        out.append(r.reshape(shape))
    return out


def response_bra_naive(j3c, j2c, mo_coeff, mo_occ, bra, scale_j, scale_k):
    """hess_r_naive.rs:861-903 (einsum form with the explicit j2c inverse).

    resp = 4 scale_j (uvP,PQ,klQ,kjA,lj,vi) - scale_k [(uvP,PQ,klQ,vjA,lj,ki) + (uvP,PQ,klQ,kjA,vj,li)]
    """
    mocc = mocc_of(mo_coeff, mo_occ)
    nao, nocc = mocc.shape
    shape = np.asarray(bra).shape
    b = np.asarray(bra).reshape(nao, nocc, -1)
    inv = np.linalg.inv(j2c)
    eri = np.einsum("uvP,PQ,klQ->uvkl", j3c, inv, j3c, optimize=True)
    rj = np.einsum("uvkl,kjA,lj,vi->uiA", eri, b, mocc, mocc, optimize=True)
    rk0 = np.einsum("uvkl,vjA,lj,ki->uiA", eri, b, mocc, mocc, optimize=True)
    rk1 = np.einsum("uvkl,kjA,vj,li->uiA", eri, b, mocc, mocc, optimize=True)
    return (4.0 * scale_j * rj - scale_k * (rk0 + rk1)).reshape(shape)


def uhf_response_bra_naive(j3c, j2c, mo_coeff, mo_occ, bra, scale_j, scale_k):
    """hess_u_naive.rs:115-167: J over both spins x2, same-spin K (two terms)."""
    inv = np.linalg.inv(j2c)
    eri = np.einsum("uvP,PQ,klQ->uvkl", j3c, inv, j3c, optimize=True)
    mocc = [mocc_of(mo_coeff[s], mo_occ[s]) for s in range(2)]
    nao = mocc[0].shape[0]
    b = [np.asarray(bra[s]).reshape(nao, mocc[s].shape[1], -1) for s in range(2)]
    out = []
    for s in range(2):
        r = np.zeros_like(b[s])
        for t in range(2):
            r += 2.0 * scale_j * np.einsum("uvkl,kjA,lj,vi->uiA", eri, b[t], mocc[t], mocc[s], optimize=True)
        r -= scale_k * np.einsum("uvkl,vjA,lj,ki->uiA", eri, b[s], mocc[s], mocc[s], optimize=True)
        r -= scale_k * np.einsum("uvkl,kjA,vj,li->uiA", eri, b[s], mocc[s], mocc[s], optimize=True)
        out.append(r.reshape(np.asarray(bra[s]).shape))
    return out


def cderi_from_j3c(j3c_tp, j2c):
    """pure_decompose.rs:107-133 + hess_r.rs:1725-1735: cderi = j3c L^-T with j2c = L L^T."""
    L = np.linalg.cholesky(j2c)
    return np.linalg.solve(L, j3c_tp.T).T


J2C_THRESH = 1e-13  # decompose.rs:5


def decomp_j2c_eig(j2c, threshold=J2C_THRESH):
    """decomp_j2c_eig (pure_decompose.rs:11-20): eigenvalues ascending, the first n = #(e < threshold)
    discarded, J^-1/2 = V_k diag(e_k^-1/2) V_k^T over the rest (symmetric)."""
    e, v = np.linalg.eigh(j2c)
    n = int((e < threshold).sum())
    vk = v[:, n:]
    return (vk * e[n:] ** -0.5) @ vk.T


def decomp_j2c_cd(j2c, threshold=None):
    """decomp_j2c_cd (pure_decompose.rs:28-55), lower factor: Cholesky; with a threshold, when the
    factorisation fails or min(diag L) <= threshold, j2c is rebuilt as V max(e, threshold) V^T and
    factorised again; with none, a failure raises (the reference panics)."""
    try:
        L = np.linalg.cholesky(j2c)
    except np.linalg.LinAlgError:
        if threshold is None:
            raise
        L = None
    if threshold is None or (L is not None and L.diagonal().min() > threshold):
        return L
    e, v = np.linalg.eigh(j2c)
    return np.linalg.cholesky((v * np.maximum(e, threshold)) @ v.T)


def cderi_from_j3c_decomp(j3c_tp, j2c, policy="cd", threshold=None):
    """generate_cderi_with_decomp (hess_r.rs:1725-1735) with solve_by_j2c side Right
    (pure_decompose.rs:179-380): cderi = j3c L^-T (policy "cd") or j3c J^-1/2 (policy "eig")."""
    if policy == "eig":
        return j3c_tp @ decomp_j2c_eig(j2c, J2C_THRESH if threshold is None else threshold)
    L = decomp_j2c_cd(j2c, threshold)
    return np.linalg.solve(L, j3c_tp.T).T
