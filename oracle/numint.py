"""Oracle: numint grid contractions for the RKS CPKS response.  Test infrastructure only.

Restates (ajz34/rstsr-showcase-hessian):
  src/numint_matmul/pure_eval_rho.rs:216-334  get_rho_from_one_bra_mult_ket_with_output
  src/numint_matmul/pure_xcpot.rs:285-352     contract_ao_wv_bra
  src/numint_matmul/pure_xcpot.rs:370-465     rks_fxc_pot_with_eff_bra_trans_with_output
  src/numint_matmul/hess_rks.rs:570-605       get_rks_response_bra (x4 closed-shell factor, 603)
Shapes follow the reference: ao [ng, nao, ncomp], rho1 [ng, nvar, nset],
fxc_eff [ng, nvar, nvar], weights [ng], out [nao, nocc, nset].
nvar: RHO 1, SIGMA 4, TAU 5 (src/xceff/flags.rs:31-38).
"""
import numpy as np

NVAR = {"rho": 1, "sigma": 4, "tau": 5}

# The reference's closed-shell / open-shell response factors, named so that the tests can plant a
# wrong one (monkeypatch) and show that the mathematical pins catch it.
RKS_RESPONSE = 4.0   # hess_rks.rs:602-603
UKS_RESPONSE = 2.0   # hess_uks.rs:436-439
NCOMP = {"rho": 1, "sigma": 4, "tau": 4}


def rho1_one_bra_mult_ket(ao, bra, kets, den):
    """pure_eval_rho.rs:216-334 (RHO/SIGMA/TAU)."""
    ng = ao.shape[0]
    nset = kets.shape[2]
    nvar = NVAR[den]
    out = np.zeros((ng, nvar, nset))
    s1 = ao[:, :, 0] @ bra  # [ng, nocc]
    for a in range(nset):
        ket = kets[:, :, a]
        s2 = ao[:, :, 0] @ ket
        out[:, 0, a] = np.sum(s1 * s2, axis=1)
        if den in ("sigma", "tau"):
            for t in range(1, 4):
                out[:, t, a] = np.sum(s1 * (ao[:, :, t] @ ket), axis=1) + np.sum((ao[:, :, t] @ bra) * s2, axis=1)
        if den == "tau":
            for t in range(1, 4):
                out[:, 4, a] += 0.5 * np.sum((ao[:, :, t] @ bra) * (ao[:, :, t] @ ket), axis=1)
    return out


def contract_ao_wv_bra(den, wv, ao, ao_bra):
    """pure_xcpot.rs:285-352: out[mu, i] for one set."""
    scr = ao_bra[:, :, 0] * wv[:, 0:1]
    out = ao[:, :, 0].T @ scr
    if den in ("sigma", "tau"):
        for t in range(1, 4):
            out += ao[:, :, 0].T @ (ao_bra[:, :, t] * wv[:, t:t + 1])
            out += ao[:, :, t].T @ (ao_bra[:, :, 0] * wv[:, t:t + 1])
    if den == "tau":
        for t in range(1, 4):
            out += 0.5 * (ao[:, :, t].T @ (ao_bra[:, :, t] * wv[:, 4:5]))
    return out


def fxc_pot_bra_trans(den, fxc_eff, rho1, ao, weights, bra, nchunk=384):
    """pure_xcpot.rs:370-465: wv[g,y,a] = w_g sum_x fxc[g,x,y] rho1[g,x,a]; per chunk x set."""
    ng, nvar, nset = rho1.shape
    nao, nocc = bra.shape
    ncomp = NCOMP[den]
    ao_bra = np.stack([ao[:, :, c] @ bra for c in range(ncomp)], axis=2)
    fw = weights[:, None, None] * fxc_eff
    out = np.zeros((nao, nocc, nset))
    for g0 in range(0, ng, nchunk):
        g1 = min(g0 + nchunk, ng)
        for a in range(nset):
            wv = np.sum(fw[g0:g1] * rho1[g0:g1, :, None, a], axis=1)
            out[:, :, a] += contract_ao_wv_bra(den, wv, ao[g0:g1], ao_bra[g0:g1])
    return out


def rks_response_bra(den, fxc_eff, ao, weights, mo1_bra, mocc):
    """hess_rks.rs:570-605: 4 * fxc_bra_trans(rho1(mocc, mo1_bra))."""
    shape = mo1_bra.shape
    nao, nocc = mocc.shape
    kets = mo1_bra.reshape(nao, nocc, -1)
    rho1 = rho1_one_bra_mult_ket(ao, mocc, kets, den)
    return (RKS_RESPONSE * fxc_pot_bra_trans(den, fxc_eff, rho1, ao, weights, mocc)).reshape(shape)


def brute_rho1(ao, bra, kets, den):
    """Per-grid-point loop form: rho1 from D = bra ket^T (independent of the GEMM ordering)."""
    ng = ao.shape[0]
    nset = kets.shape[2]
    out = np.zeros((ng, NVAR[den], nset))
    for a in range(nset):
        D = bra @ kets[:, :, a].T
        for g in range(ng):
            p = ao[g]
            out[g, 0, a] = p[:, 0] @ D @ p[:, 0]
            if den in ("sigma", "tau"):
                for t in range(1, 4):
                    out[g, t, a] = p[:, 0] @ D @ p[:, t] + p[:, t] @ D @ p[:, 0]
            if den == "tau":
                out[g, 4, a] = 0.5 * sum(p[:, t] @ D @ p[:, t] for t in range(1, 4))
    return out


def brute_fxc_full(den, fxc_eff, rho1, ao, weights):
    """Full AO-pair potential V[mu, nu, a] by a per-grid-point loop (symmetric in mu <-> nu after
    adding its transpose is NOT applied: matches the non-symmetrised bra form when right-multiplied)."""
    ng, nvar, nset = rho1.shape
    nao = ao.shape[1]
    V = np.zeros((nao, nao, nset))
    for a in range(nset):
        for g in range(ng):
            wv = weights[g] * (fxc_eff[g].T @ rho1[g, :, a])
            p = ao[g]
            V[:, :, a] += wv[0] * np.outer(p[:, 0], p[:, 0])
            if den in ("sigma", "tau"):
                for t in range(1, 4):
                    V[:, :, a] += wv[t] * (np.outer(p[:, 0], p[:, t]) + np.outer(p[:, t], p[:, 0]))
            if den == "tau":
                for t in range(1, 4):
                    V[:, :, a] += 0.5 * wv[4] * np.outer(p[:, t], p[:, t])
    return V


# ---------------------------------------------------------------------------------------------
# UKS (pure_xcpot.rs:921-1028 uks_fxc_pot_with_eff_bra_trans_with_output; hess_uks.rs:390-441
# get_uks_response_bra, factor 2).  fxc_eff [ng, nvar, 2, nvar, 2], rho1 [ng, nvar, 2, nset].

def uks_fxc_pot_bra_trans(den, fxc_eff, rho1, ao, weights, bra):
    """wv_s[g, x, a] = w_g sum_{y,t} fxc[g, y, t, x, s] rho1[g, y, t, a]; per spin bra contraction."""
    ncomp = NCOMP[den]
    nset = rho1.shape[3]
    out = []
    for s in range(2):
        ao_bra = np.stack([ao[:, :, c] @ bra[s] for c in range(ncomp)], axis=2)
        o = np.zeros((ao.shape[1], bra[s].shape[1], nset))
        for a in range(nset):
            wv = weights[:, None] * np.einsum("gytx,gyt->gx", fxc_eff[:, :, :, :, s], rho1[:, :, :, a])
            o[:, :, a] = contract_ao_wv_bra(den, wv, ao, ao_bra)
        out.append(o)
    return out


def uks_response_bra(den, fxc_eff, ao, weights, mo1_bra, mocc):
    """hess_uks.rs:390-441: rho1 per spin, stacked, then 2 * uks_fxc_pot_bra_trans."""
    kets = [m.reshape(m.shape[0], m.shape[1], -1) for m in mo1_bra]
    rho1 = np.stack([rho1_one_bra_mult_ket(ao, mocc[s], kets[s], den) for s in range(2)], axis=2)
    r = uks_fxc_pot_bra_trans(den, fxc_eff, rho1, ao, weights, mocc)
    return [UKS_RESPONSE * r[s].reshape(mo1_bra[s].shape) for s in range(2)]


def brute_uks_potential(den, fxc_eff, rho1, ao, weights):
    """Per-grid-point AO-pair potentials V_s[mu, nu, a] with the spin-coupled kernel."""
    ng, nvar, _, nset = rho1.shape
    nao = ao.shape[1]
    V = [np.zeros((nao, nao, nset)) for _ in range(2)]
    for s in range(2):
        for a in range(nset):
            for g in range(ng):
                wv = weights[g] * np.einsum("ytx,yt->x", fxc_eff[g, :, :, :, s], rho1[g, :, :, a])
                p = ao[g]
                V[s][:, :, a] += wv[0] * np.outer(p[:, 0], p[:, 0])
                if den in ("sigma", "tau"):
                    for t in range(1, 4):
                        V[s][:, :, a] += wv[t] * (np.outer(p[:, 0], p[:, t]) + np.outer(p[:, t], p[:, 0]))
                if den == "tau":
                    for t in range(1, 4):
                        V[s][:, :, a] += 0.5 * wv[4] * np.outer(p[:, t], p[:, t])
    return V
