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
