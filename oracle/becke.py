"""Oracle: Becke partition weights and their nuclear derivatives (NumPy FP64).  Test
infrastructure only.

Restates (ajz34/rstsr-showcase-hessian) src/numint_matmul/becke_partition.rs:
  163-205    BeckeMolTables::new (interatomic distances, their first derivatives)
  232-340    becke_partition / becke_partition_with_tables (ByGrid / ByAtom attribution,
             outputs w, dw, ddw and the contractions c, dc, ddc)
  796-863    eval_partition (distances, switch products P_A, Z, w = wquad P_g / Z)
  865-1100   eval_switch_pair_pass (first role derivatives, dR_Z, dR_Pg; the per-pair
             second-order terms of dR_log_P, ddR_Z, ddR_Pg)
  1102-1153  eval_lane_dw (quotient rule, translation-invariance fix of the generating atom's row)
  1155-1380  eval_lane_ddw (rank-1 cross terms, quotient rule, row / column / corner fix)
  1513-1590  switch functions, general hardness recurrences
The grid point belongs to its generating atom g(r): dw / ddw are total derivatives with the grid
carried along by that atom (the reference's translation-invariance fix).  Becke, J. Chem. Phys.
88, 2547 (1988), eqs. (A2), (19), (20).

Layouts: grid_coords [ng, 3]; atm_coords [natm, 3]; attribution [ng] (atom index; >= natm is
padding and carries no weight); adjustment [natm, natm] row-major (entry (A, B) is a_AB);
outputs w [ng], dw [natm, 3, ng], ddw [natm, 3, natm, 3, ng], c [nset], dc [natm, 3, nset],
ddc [natm, 3, natm, 3, nset] (C order, as the reference's flat buffers).
"""
import numpy as np

INVTOL = 1e-14  # becke_partition.rs:24


def mol_tables(atm_coords):
    """becke_partition.rs:163-205: distances (inf on the diagonal) and (R_A - R_B) / |R_AB|."""
    R = np.asarray(atm_coords, dtype=np.float64)
    diff = R[:, None, :] - R[None, :, :]
    dist = np.linalg.norm(diff, axis=-1)
    np.fill_diagonal(dist, np.inf)
    return dist, diff / dist[:, :, None]


def switch(mu, a, hardness):
    """becke_partition.rs:1575-1590: f, df/dnu, d2f/dnu2 of f = p^hardness(nu), p(x) = 3x/2 - x^3/2,
    nu = mu + a (1 - mu^2)."""
    nu = mu + (1.0 - mu * mu) * a
    f, df, ddf = nu, np.ones_like(nu), np.zeros_like(nu)
    for _ in range(hardness):
        g = 1.5 * (1.0 - f * f)
        ddf = -3.0 * f * df * df + g * ddf
        df = g * df
        f = (1.5 - 0.5 * f * f) * f
    return f, df, ddf


def _lane_batch(R, atm_dist, dR_atm, adj, coords, wquad, attr, hardness, deriv):
    """One batch of grid points (the reference's lanes), grid axis first."""
    n, natm = coords.shape[0], R.shape[0]
    n3 = 3 * natm
    rows = np.arange(n)
    valid = attr < natm
    ag = np.where(valid, attr, 0)

    # --- deriv 0 (eval_partition, becke_partition.rs:796-863) ---
    dist = np.linalg.norm(coords[:, None, :] - R[None, :, :], axis=-1)       # [n, natm]
    P = np.ones((n, natm))
    for A in range(natm):
        for B in range(A):
            mu = (dist[:, A] - dist[:, B]) / atm_dist[A, B]
            f = switch(mu, adj[A, B], hardness)[0]
            P[:, A] *= 0.5 * (1.0 - f)
            P[:, B] *= 0.5 * (1.0 + f)
    Z = P.sum(axis=1)
    Pg = np.where(valid, P[rows, ag], 0.0)
    out = dict(w=wquad * Pg / Z)
    if deriv < 1:
        return out

    # --- first role derivatives and the pair pass (becke_partition.rs:865-1100) ---
    dR_dist = (R[None, :, :] - coords[:, None, :]) / dist[:, :, None]       # [n, natm, 3]
    dR_Z = np.zeros((n, natm, 3))
    dR_Pg = np.zeros((n, natm, 3))
    if deriv >= 2:
        dlogP = np.zeros((n, natm, natm, 3))      # [g, M, C, t]
        ddZ = np.zeros((n, natm, 3, natm, 3))
        ddPg = np.zeros((n, natm, 3, natm, 3))
        eye = np.eye(3)
        PrM = (eye[None, None] - dR_dist[:, :, :, None] * dR_dist[:, :, None, :]) / dist[:, :, None, None]
    for A in range(natm):
        for B in range(A):
            a, gab = adj[A, B], atm_dist[A, B]
            mu = (dist[:, A] - dist[:, B]) / gab
            f, df, ddf = switch(mu, a, hardness)
            sA, sB = 0.5 * (1.0 - f), 0.5 * (1.0 + f)
            dmu_nu = 1.0 - 2.0 * mu * a
            dsA, dsB = -0.5 * df * dmu_nu, 0.5 * df * dmu_nu
            sA_safe, sB_safe = np.maximum(sA, INVTOL), np.maximum(sB, INVTOL)
            lA, lB = dsA / sA_safe, dsB / sB_safe
            cZ = P[:, A] * lA + P[:, B] * lB
            selA = valid & (attr == A)
            selB = valid & (attr == B)
            cPg = np.where(selA, P[:, A] * lA, 0.0) + np.where(selB, P[:, B] * lB, 0.0)
            U = dR_atm[A, B]
            roleA = (dR_dist[:, A] - mu[:, None] * U) / gab
            roleB = (-dR_dist[:, B] + mu[:, None] * U) / gab
            dR_Z[:, A] += cZ[:, None] * roleA
            dR_Z[:, B] += cZ[:, None] * roleB
            dR_Pg[:, A] += cPg[:, None] * roleA
            dR_Pg[:, B] += cPg[:, None] * roleB
            if deriv < 2:
                continue
            ddmu_nu = -2.0 * a
            ddsA = -0.5 * ddf * dmu_nu * dmu_nu - 0.5 * df * ddmu_nu
            ddlA = ddsA / sA_safe - lA * lA
            ddlB = -ddsA / sB_safe - lB * lB
            # second role derivatives of mu_AB by the quotient rule (becke_partition.rs:1008-1041)
            rA, rB = dR_dist[:, A], dR_dist[:, B]
            PU = (eye - np.outer(U, U)) / gab
            fab = dist[:, A] - dist[:, B]

            def d2mu(fX, fY, fXY, gX, gY, gXY):
                ofg = fX[:, :, None] * gY[None, None, :] + gX[None, :, None] * fY[:, None, :]
                ogg = np.outer(gX, gY)[None]
                return ((fXY * gab - ofg - fab[:, None, None] * gXY[None]) / gab ** 2
                        + 2.0 * fab[:, None, None] * ogg / gab ** 3)

            hAA = d2mu(rA, rA, PrM[:, A], U, U, PU)
            hAB = d2mu(rA, -rB, np.zeros((n, 3, 3)), U, -U, -PU)
            hBB = d2mu(-rB, -rB, -PrM[:, B], -U, -U, PU)
            dlogP[:, A, A] += lA[:, None] * roleA
            dlogP[:, A, B] = lA[:, None] * roleB
            dlogP[:, B, A] = lB[:, None] * roleA
            dlogP[:, B, B] += lB[:, None] * roleB
            cdd = P[:, A] * ddlA + P[:, B] * ddlB
            cA = np.where(selA, P[:, A], 0.0)
            cB = np.where(selB, P[:, B], 0.0)
            c1g = cA * lA + cB * lB
            cddg = cA * ddlA + cB * ddlB
            oo = lambda x, y: x[:, :, None] * y[:, None, :]
            for buf, c1, cd in ((ddZ, cZ, cdd), (ddPg, c1g, cddg)):
                buf[:, A, :, A, :] += cd[:, None, None] * oo(roleA, roleA) + c1[:, None, None] * hAA
                buf[:, A, :, B, :] = cd[:, None, None] * oo(roleA, roleB) + c1[:, None, None] * hAB
                buf[:, B, :, A, :] = cd[:, None, None] * oo(roleB, roleA) + c1[:, None, None] * hAB.transpose(0, 2, 1)
                buf[:, B, :, B, :] += cd[:, None, None] * oo(roleB, roleB) + c1[:, None, None] * hBB

    # --- dw: quotient rule and the generating atom's row (becke_partition.rs:1102-1153) ---
    dZ = dR_Z.reshape(n, n3)
    dPg = dR_Pg.reshape(n, n3)
    dw = (wquad / Z)[:, None] * (dPg - (Pg / Z)[:, None] * dZ)
    dw = dw.reshape(n, natm, 3)
    neg = -dw.sum(axis=1)
    for i in np.nonzero(valid)[0]:
        dw[i, ag[i]] += neg[i]
    out["dw"] = dw
    if deriv < 2:
        return out

    # --- ddw (becke_partition.rs:1155-1380) ---
    L = dlogP.reshape(n, natm, n3)
    C = np.einsum("nm,nma,nmb->nab", P, L, L)
    LG = L[rows, ag] * valid[:, None]
    PG = np.where(valid, P[rows, ag], 0.0)
    q = Pg / Z
    dq = (dPg - q[:, None] * dZ) / Z[:, None]
    ddZf = ddZ.reshape(n, n3, n3) + C
    ddPf = ddPg.reshape(n, n3, n3) + PG[:, None, None] * LG[:, :, None] * LG[:, None, :]
    d2q = ((ddPf - dZ[:, :, None] * dq[:, None, :] - q[:, None, None] * ddZf) / Z[:, None, None]
           - dq[:, :, None] * dZ[:, None, :] / Z[:, None, None])
    v = (wquad[:, None, None] * d2q).reshape(n, natm, 3, natm, 3)
    ddw = v.copy()
    for i in np.nonzero(valid)[0]:
        G = ag[i]
        fullA = v[i].sum(axis=0)           # [t, B, s]
        fullB = v[i].sum(axis=2)           # [A, t, s]
        fullAB = fullA.sum(axis=1)         # [t, s]
        # row A = G: -sum_{A' != G}; corner: sum over both other axes
        ddw[i, G] = -fullA + v[i, G]
        ddw[i, G, :, G] = fullAB - fullB[G] - fullA[:, G] + v[i, G, :, G]
        for A in range(natm):
            if A != G:
                ddw[i, A, :, G] -= fullB[A]
    out["ddw"] = ddw
    return out


def partition(grid_coords, atm_coords, attribution, wquad, adjustment, hardness=3, deriv=2,
              contract_w=None, contract_dw=None, contract_ddw=None, output=(True, True, True),
              nbatch=256):
    """becke_partition.rs:232-340.  attribution: per-grid atom indices (ByGrid).  contract_*:
    [nset, ng] weights of the c / dc / ddc contractions.  Returns a dict with the requested
    outputs (w, dw, ddw in the layouts of the module docstring, and c, dc, ddc)."""
    coords = np.asarray(grid_coords, dtype=np.float64)
    R = np.asarray(atm_coords, dtype=np.float64)
    attr = np.asarray(attribution, dtype=np.int64)
    wquad = np.asarray(wquad, dtype=np.float64)
    adj = np.asarray(adjustment, dtype=np.float64)
    ng, natm = coords.shape[0], R.shape[0]
    atm_dist, dR_atm = mol_tables(R)
    cw = [None if c is None else np.atleast_2d(np.asarray(c, dtype=np.float64)) for c in
          (contract_w, contract_dw, contract_ddw)]
    need2 = deriv >= 2 and (output[2] or cw[2] is not None)
    d_eval = min(deriv, 2 if need2 else 1)
    res = dict(w=np.zeros(ng))
    if deriv >= 1:
        res["dw"] = np.zeros((natm, 3, ng))
    if deriv >= 2 and output[2]:
        res["ddw"] = np.zeros((natm, 3, natm, 3, ng))
    if cw[0] is not None:
        res["c"] = np.zeros(cw[0].shape[0])
    if deriv >= 1 and cw[1] is not None:
        res["dc"] = np.zeros((natm, 3, cw[1].shape[0]))
    if need2 and cw[2] is not None:
        res["ddc"] = np.zeros((natm, 3, natm, 3, cw[2].shape[0]))
    for g0 in range(0, ng, nbatch):
        s = slice(g0, min(ng, g0 + nbatch))
        o = _lane_batch(R, atm_dist, dR_atm, adj, coords[s], wquad[s], attr[s], hardness, d_eval)
        res["w"][s] = o["w"]
        if cw[0] is not None:
            res["c"] += cw[0][:, s] @ o["w"]
        if deriv >= 1:
            dwb = o["dw"].transpose(1, 2, 0)             # [A, t, g]
            res["dw"][:, :, s] = dwb
            if cw[1] is not None:
                res["dc"] += np.einsum("atg,kg->atk", dwb, cw[1][:, s])
        if need2:
            ddb = o["ddw"].transpose(1, 2, 3, 4, 0)      # [A, t, B, s, g]
            if output[2]:
                res["ddw"][..., s] = ddb
            if cw[2] is not None:
                res["ddc"] += np.einsum("atbsg,kg->atbsk", ddb, cw[2][:, s])
    keep = {"w": output[0], "dw": output[1] and deriv >= 1, "ddw": output[2] and deriv >= 2}
    return {k: v for k, v in res.items() if keep.get(k, True)}


def by_atom_attribution(split):
    """ByAtom boundaries [natm + 1] -> per-grid attribution (becke_partition.rs:125-141)."""
    split = np.asarray(split, dtype=np.int64)
    natm = len(split) - 1
    if np.any(np.diff(split) < 0) or split[0] != 0:
        raise ValueError("ByAtom boundaries must be non-decreasing from 0")
    return np.repeat(np.arange(natm), np.diff(split))


def partition_weights_loop(grid_coords, atm_coords, attribution, wquad, adjustment, hardness):
    """Zeroth order by an explicit pair loop with the switch iterated directly (no derivative
    machinery), for the general-hardness check of w."""
    R = np.asarray(atm_coords, dtype=np.float64)
    coords = np.asarray(grid_coords, dtype=np.float64)
    natm = R.shape[0]
    out = np.zeros(coords.shape[0])
    for i, r in enumerate(coords):
        d = np.linalg.norm(r - R, axis=1)
        P = np.ones(natm)
        for A in range(natm):
            for B in range(natm):
                if A == B:
                    continue
                mu = (d[A] - d[B]) / np.linalg.norm(R[A] - R[B])
                nu = mu + adjustment[A][B] * (1.0 - mu * mu)
                for _ in range(hardness):
                    nu = 1.5 * nu - 0.5 * nu ** 3
                P[A] *= 0.5 * (1.0 - nu)
        if attribution[i] < natm:
            out[i] = wquad[i] * P[attribution[i]] / P.sum()
    return out
