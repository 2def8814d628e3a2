"""Oracle: LibXC-variable derivatives -> rho_t-based effective kernels (unpolarised).

Test infrastructure only.  Restates src/xceff/xc_deriv.rs (ajz34/rstsr-showcase-hessian):
  get_xc_xlen            97-107   (SIGMA: xlen 2 (rho, sigma); TAU: xlen 3 (rho, sigma, tau))
  product_uniq_indices   111-134  (unique derivative columns = combinations with replacement)
  vxc_unfold_sigma_spin0 140-186  (f_rho_t = 2 f_sigma grad_t rho)
  transform_xc_inner     305-428  (order 2 adds 2 f_sigma on the gradient diagonal, 362-376)
xc_val columns are [exc | 1st derivs (xlen) | 2nd derivs (xlen(xlen+1)/2)], offsets as in
transform_xc_inner (325-327).
"""
import itertools

import numpy as np

XLEN = {"sigma": 2, "tau": 3}
NVAR = {"sigma": 4, "tau": 5}


def uniq_pairs(xlen):
    return list(itertools.combinations_with_replacement(range(xlen), 2))


def transform(rho, xc_val, den, order):
    """Returns vxc_eff [ng, nvar] (order 1) or fxc_eff [ng, nvar, nvar] (order 2)."""
    xlen, nvar = XLEN[den], NVAR[den]
    ng = rho.shape[0]
    # which LibXC variable each rho_t component maps to, and its chain-rule coefficient
    var = [0, 1, 1, 1] + ([2] if den == "tau" else [])
    coef = np.ones((ng, nvar))
    coef[:, 1:4] = 2.0 * rho[:, 1:4]
    v1 = xc_val[:, 1:1 + xlen]
    if order == 1:
        return v1[:, var] * coef
    assert order == 2
    f2 = np.zeros((ng, xlen, xlen))
    for k, (i, j) in enumerate(uniq_pairs(xlen)):
        f2[:, i, j] = f2[:, j, i] = xc_val[:, 1 + xlen + k]
    out = f2[:, var][:, :, var] * coef[:, :, None] * coef[:, None, :]
    for t in range(1, 4):
        out[:, t, t] += 2.0 * v1[:, 1]
    return out


# ---------------------------------------------------------------------------------------------
# Spin-polarised transform (xc_deriv.rs:188-243 vxc_unfold_sigma_spin1, 379-424 order-2 sigma
# pair term) and the LibXC -> XCFun column order (xc_deriv.rs:17-88).
#
# XCFun variable order (spin polarised): rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb[, tau_a,
# tau_b] (get_xc_xlen 97-107: RHO 2, SIGMA 5, TAU 7).  rho [ng, nvar, 2] (x = rho, grad_x,
# grad_y, grad_z, tau; last index the spin); outputs vxc_eff [ng, nvar, 2] and fxc_eff
# [ng, nvar, 2, nvar, 2].

XLEN_POL = {"rho": 2, "sigma": 5, "tau": 7}
NVAR_POL = {"rho": 1, "sigma": 4, "tau": 5}


def jacobian_pol(rho, den):
    """J[g, k, x, s] = d(XCFun variable k) / d(rho_t variable x of spin s)."""
    ng = rho.shape[0]
    xlen, nvar = XLEN_POL[den], NVAR_POL[den]
    J = np.zeros((ng, xlen, nvar, 2))
    J[:, 0, 0, 0] = 1.0
    J[:, 1, 0, 1] = 1.0
    if den != "rho":
        for t in range(1, 4):
            J[:, 2, t, 0] = 2.0 * rho[:, t, 0]           # sigma_aa = grad a . grad a
            J[:, 3, t, 0] = rho[:, t, 1]                 # sigma_ab = grad a . grad b
            J[:, 3, t, 1] = rho[:, t, 0]
            J[:, 4, t, 1] = 2.0 * rho[:, t, 1]           # sigma_bb
    if den == "tau":
        J[:, 5, 4, 0] = 1.0
        J[:, 6, 4, 1] = 1.0
    return J


def transform_pol(rho, xc_val, den, order):
    """transform_xc_inner with spin = Polarized for order 1 and 2; xc_val in XCFun order."""
    xlen, nvar = XLEN_POL[den], NVAR_POL[den]
    ng = rho.shape[0]
    J = jacobian_pol(rho, den)
    v1 = xc_val[:, 1:1 + xlen]
    if order == 1:
        return np.einsum("gk,gkxs->gxs", v1, J)
    assert order == 2
    f2 = np.zeros((ng, xlen, xlen))
    for k, (i, j) in enumerate(uniq_pairs(xlen)):
        f2[:, i, j] = f2[:, j, i] = xc_val[:, 1 + xlen + k]
    out = np.einsum("gkxs,gkl,glyt->gxsyt", J, f2, J)
    if den != "rho":
        # second derivative of the sigma variables in grad rho: [[2 f_aa, f_ab], [f_ab, 2 f_bb]]
        M = np.stack([np.stack([2.0 * v1[:, 2], v1[:, 3]], -1), np.stack([v1[:, 3], 2.0 * v1[:, 4]], -1)], -2)
        for t in range(1, 4):
            out[:, t, :, t, :] += M
    return out


def libxc_to_xcfun_map(den, deriv):
    """Columns of the combined LibXC output [zk | 1st | 2nd] listed in XCFun order, polarised,
    deriv <= 2.  LibXC groups its polarised derivatives by variable kind (rho 2, sigma 3, tau 2
    components) and its second-derivative blocks by kind pairs in the order (rho,rho),
    (rho,sigma), (rho,tau), (sigma,sigma), (sigma,tau), (tau,tau), components row-major inside a
    block (upper triangle for a kind with itself); XCFun lists all pairs of the flat variable
    list as combinations with replacement."""
    kinds = {"rho": [2], "sigma": [2, 3], "tau": [2, 3, 2]}[den]
    flat = [(ki, c) for ki, n in enumerate(kinds) for c in range(n)]
    out = [0]
    if deriv >= 1:
        out += list(range(1, 1 + len(flat)))
    if deriv >= 2:
        off = 1 + len(flat)
        pos = {}
        for k1, k2 in itertools.combinations_with_replacement(range(len(kinds)), 2):
            if k1 == k2:
                pairs = itertools.combinations_with_replacement(range(kinds[k1]), 2)
            else:
                pairs = itertools.product(range(kinds[k1]), range(kinds[k2]))
            for c1, c2 in pairs:
                pos[((k1, c1), (k2, c2))] = off
                off += 1
        for i, j in itertools.combinations_with_replacement(range(len(flat)), 2):
            out.append(pos[(flat[i], flat[j])])
    if deriv >= 3:
        raise NotImplementedError("orders above 2 are not needed for Hessians")
    return out
