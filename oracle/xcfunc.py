"""Oracle: exchange-correlation functionals with first and second derivatives.

Test infrastructure only (never shipped, never on the product path).  The reference obtains
these values from LibXC through the `libxc` crate (Cargo.toml:16, `libxc = "0.1"`, a binding
of the system LibXC; neither is in the reference tree):
  libxc_eval_inner   src/xceff/libxc_wrap.rs:32-82   xc_func.compute_xc(rho, sigma; deriv)
  libxc_eval_eff     src/xceff/libxc_wrap.rs:88-122  + libxc_transform_xcfun_indices, transform_xc_inner
  functional lists   tests/hessian_rks/test_svwn.rs:28-29 (LDA_X + LDA_C_VWN),
                     tests/hessian_rks/test_b3lyp.rs:14 (HYB_GGA_XC_B3LYP), UKS tests the same, polarised.
The functionals are restated from their publications (LibXC implements the same formulas):
  Slater exchange; Becke 1988 exchange (beta = 0.0042); Vosko-Wilk-Nusair 1980 correlation with
  the fit-V (VWN5) and RPA parameter sets and the spin interpolation through the spin stiffness;
  Lee-Yang-Parr correlation in the Laplacian-free form of Miehlich, Savin, Stoll, Preuss 1989;
  B3LYP = 0.08 Slater + 0.72 B88 + 0.19 VWN_RPA + 0.81 LYP (+ 0.2 exact exchange).
Derivatives come from torch.autograd (float64).  Output columns follow LibXC: unpolarised
[zk, vrho, vsigma, v2rho2, v2rhosigma, v2sigma2]; polarised [zk, vrho(a, b), vsigma(aa, ab, bb),
v2rho2(aa, ab, bb), v2rhosigma(a x sigma, b x sigma), v2sigma2(upper triangle, row-major)].

Pins (tests/test_oracle_xcfunc.py): Slater's closed form and exchange spin scaling; B88's LDA
limit, gradient-expansion coefficient and large-gradient asymptote (independent of beta); VWN's
high-density logarithm and constant (exact 0.0311 ln rs - 0.047 for the fit, RPA constant
-0.071), agreement of the fit-V curves with the independent Perdew-Wang 1992 parametrisation of
the same Monte Carlo data, and the spin-stiffness curvature in zeta; LYP's reduction to the
published closed-shell form and its vanishing for fully polarised densities; the unpolarised
functional as the polarised one at equal spins; autograd derivatives against finite differences.
Parity unpinned: which spin interpolation LibXC's LDA_C_VWN_RPA uses for zeta != 0 (restated here
as VWN5's, with the RPA parameters, the LibXC description "VWN5_RPA"); unpolarised values are
unaffected.
"""
import math

import numpy as np
import torch

IDS = {"LDA_X": 1, "LDA_C_VWN": 7, "LDA_C_VWN_RPA": 8, "GGA_X_B88": 106, "GGA_C_LYP": 131, "HYB_GGA_XC_B3LYP": 402}
GGA = {1: False, 7: False, 8: False, 106: True, 131: True, 402: True}
EXX = {402: 0.2}

DENS_THR = 1e-15
ZETA_THR = 2.220446049250313e-16
SIGMA_FLOOR = 1e-40

CX_SPIN = 0.75 * (6.0 / math.pi) ** (1.0 / 3.0)   # exchange per spin: -CX_SPIN rho_s^(4/3)
B88_BETA = 0.0042

# VWN parameters (A, b, c, x0): paramagnetic, ferromagnetic, spin stiffness
VWN5 = [(0.0310907, 3.72744, 12.9352, -0.10498),
        (0.01554535, 7.06042, 18.0578, -0.32500),
        (-1.0 / (6.0 * math.pi ** 2), 1.13107, 13.0045, -0.0047584)]
VWN_RPA = [(0.0310907, 13.0720, 42.7198, -0.409286),
           (0.01554535, 20.1231, 101.578, -0.743294),
           (-1.0 / (6.0 * math.pi ** 2), 1.06835, 11.4813, -0.228344)]

LYP_A, LYP_B, LYP_C, LYP_D = 0.04918, 0.132, 0.2533, 0.349
CF = 0.3 * (3.0 * math.pi ** 2) ** (2.0 / 3.0)


def vwn_curve(p, rs):
    """VWN interpolation F(rs) of one curve: A [ln(x^2/X) + 2b/Q atan(Q/(2x+b))
    - b x0/X(x0) (ln((x-x0)^2/X) + 2(b+2x0)/Q atan(Q/(2x+b)))], x = sqrt(rs)."""
    A, b, c, x0 = p
    x = torch.sqrt(rs)
    X = rs + b * x + c
    Q = math.sqrt(4.0 * c - b * b)
    X0 = x0 * x0 + b * x0 + c
    at = torch.atan(Q / (2.0 * x + b))
    return A * (torch.log(rs / X) + 2.0 * b / Q * at
                - b * x0 / X0 * (torch.log((x - x0) ** 2 / X) + 2.0 * (b + 2.0 * x0) / Q * at))


def f_zeta(z):
    """Spin interpolation f(zeta) = ((1+z)^(4/3) + (1-z)^(4/3) - 2) / (2^(4/3) - 2), with each
    1 +- z floored at the zeta threshold."""
    def p43(t):
        return torch.where(t <= ZETA_THR, torch.full_like(t, ZETA_THR ** (4.0 / 3.0)),
                           torch.clamp(t, min=ZETA_THR) ** (4.0 / 3.0))
    return (p43(1.0 + z) + p43(1.0 - z) - 2.0) / (2.0 ** (4.0 / 3.0) - 2.0)


FPP = 4.0 / (9.0 * (2.0 ** (1.0 / 3.0) - 1.0))   # f''(0)


def vwn_eps(params, rs, z):
    eP, eF, eA = (vwn_curve(p, rs) for p in params)
    fz = f_zeta(z)
    z4 = z ** 4
    return eP + eA * fz * (1.0 - z4) / FPP + (eF - eP) * fz * z4


def vwn(params, ra, rb):
    rho = ra + rb
    rs = (3.0 / (4.0 * math.pi * rho)) ** (1.0 / 3.0)
    return rho * vwn_eps(params, rs, (ra - rb) / rho)


def slater_spin(r):
    return -CX_SPIN * r ** (4.0 / 3.0)


def b88_spin(r, s):
    """Becke 88 exchange of one spin channel: Slater - beta r^(4/3) x^2 / (1 + 6 beta x asinh x),
    x = sqrt(s) / r^(4/3)."""
    x = torch.sqrt(s) / r ** (4.0 / 3.0)
    return slater_spin(r) - B88_BETA * r ** (4.0 / 3.0) * x * x / (1.0 + 6.0 * B88_BETA * x * torch.asinh(x))


def lyp(ra, rb, saa, sab, sbb):
    """LYP, Miehlich et al. 1989 (no Laplacian)."""
    a, b, c, d = LYP_A, LYP_B, LYP_C, LYP_D
    rho = ra + rb
    rm = rho ** (-1.0 / 3.0)
    den = 1.0 + d * rm
    omega = torch.exp(-c * rm) / den * rho ** (-11.0 / 3.0)
    delta = c * rm + d * rm / den
    g = saa + 2.0 * sab + sbb
    br = (2.0 ** (11.0 / 3.0) * CF * (ra ** (8.0 / 3.0) + rb ** (8.0 / 3.0))
          + (47.0 / 18.0 - 7.0 / 18.0 * delta) * g
          - (2.5 - delta / 18.0) * (saa + sbb)
          - (delta - 11.0) / 9.0 * (ra / rho * saa + rb / rho * sbb))
    brace = (ra * rb * br - 2.0 / 3.0 * rho ** 2 * g
             + (2.0 / 3.0 * rho ** 2 - ra ** 2) * sbb + (2.0 / 3.0 * rho ** 2 - rb ** 2) * saa)
    return -4.0 * a / den * ra * rb / rho - a * b * omega * brace


def energy_density(xc_id, ra, rb, saa, sab, sbb, on_a, on_b):
    """Energy per volume of functional xc_id over the polarised variables; on_s masks the
    exchange of a spin channel at or below the density threshold."""
    def slater():
        return torch.where(on_a, slater_spin(ra), 0.0 * ra) + torch.where(on_b, slater_spin(rb), 0.0 * rb)

    def b88():
        return torch.where(on_a, b88_spin(ra, saa), 0.0 * ra) + torch.where(on_b, b88_spin(rb, sbb), 0.0 * rb)

    if xc_id == 1:
        return slater()
    if xc_id == 7:
        return vwn(VWN5, ra, rb)
    if xc_id == 8:
        return vwn(VWN_RPA, ra, rb)
    if xc_id == 106:
        return b88()
    if xc_id == 131:
        return lyp(ra, rb, saa, sab, sbb)
    if xc_id == 402:
        return 0.08 * slater() + 0.72 * b88() + 0.19 * vwn(VWN_RPA, ra, rb) + 0.81 * lyp(ra, rb, saa, sab, sbb)
    raise ValueError(f"unsupported functional id {xc_id}")


def _variables(rho, nspin, gga):
    """LibXC variables from rho [ng, nvar] (nspin 1) or [ng, nvar, 2] (nspin 2)."""
    rho = torch.as_tensor(np.asarray(rho), dtype=torch.float64)
    if nspin == 1:
        v = [rho[:, 0].clone()]
        if gga:
            v.append(torch.clamp((rho[:, 1:4] ** 2).sum(1), min=SIGMA_FLOOR))
        rtot = rho[:, 0]
    else:
        v = [rho[:, 0, 0].clone(), rho[:, 0, 1].clone()]
        if gga:
            ga, gb = rho[:, 1:4, 0], rho[:, 1:4, 1]
            v += [torch.clamp((ga * ga).sum(1), min=SIGMA_FLOOR), (ga * gb).sum(1),
                  torch.clamp((gb * gb).sum(1), min=SIGMA_FLOOR)]
        rtot = rho[:, 0, 0] + rho[:, 0, 1]
    return v, rtot


def _polarised_args(v, nspin, gga):
    if nspin == 1:
        r = v[0]
        ra = rb = 0.5 * r
        s = 0.25 * v[1] if gga else torch.zeros_like(r)
        on = torch.ones_like(r, dtype=torch.bool)
        return ra, rb, s, s, s, on, on
    on_a, on_b = v[0] > DENS_THR, v[1] > DENS_THR
    ra = torch.clamp(v[0], min=DENS_THR)
    rb = torch.clamp(v[1], min=DENS_THR)
    if gga:
        return ra, rb, v[2], v[3], v[4], on_a, on_b
    z = torch.zeros_like(ra)
    return ra, rb, z, z, z, on_a, on_b


def libxc_order_pairs(nvar, nspin, gga):
    """Second-derivative variable pairs in LibXC's column order."""
    if nspin == 1 or not gga:
        return [(i, j) for i in range(nvar) for j in range(i, nvar)]
    return ([(0, 0), (0, 1), (1, 1)] + [(i, j) for i in range(2) for j in range(2, 5)]
            + [(i, j) for i in range(2, 5) for j in range(i, 5)])


def eval_raw(xc_id, rho, nspin, deriv):
    """LibXC output [ng, ncol] (numpy) of functional xc_id at rho."""
    gga = GGA[xc_id]
    v, rtot = _variables(rho, nspin, gga)
    for t in v:
        t.requires_grad_(True)
    e = energy_density(xc_id, *_polarised_args(v, nspin, gga))
    ok = rtot >= DENS_THR
    cols = [(e / rtot).detach()]
    if deriv >= 1:
        g = torch.autograd.grad(e.sum(), v, create_graph=deriv >= 2, allow_unused=True)
        g = [gi if gi is not None else torch.zeros_like(v[0]) for gi in g]
        cols += [gi.detach() for gi in g]
        if deriv >= 2:
            h = {}
            for i in range(len(v)):
                if not g[i].requires_grad:
                    hi = [None] * len(v)
                else:
                    hi = torch.autograd.grad(g[i].sum(), v, retain_graph=True, allow_unused=True)
                for j in range(len(v)):
                    h[i, j] = hi[j].detach() if hi[j] is not None else torch.zeros_like(v[0])
            cols += [h[i, j] for i, j in libxc_order_pairs(len(v), nspin, gga)]
    out = torch.stack(cols, 1)
    out[~ok] = 0.0
    return out.numpy()


def ncol(xc_id, nspin, deriv):
    n = (2 if GGA[xc_id] else 1) if nspin == 1 else (5 if GGA[xc_id] else 2)
    return 1 + (n if deriv >= 1 else 0) + (n * (n + 1) // 2 if deriv >= 2 else 0)


def eval_eff(xc_id, rho, nspin, den_list):
    """libxc_eval_eff at order 2 embedded in the leading slice of the caller's arrays:
    (exc [ng], vxc [ng, nvl(, 2)], fxc [ng, nvl(, 2), nvl(, 2)]) with nvl that of den_list.
    Chain rule by the separately pinned transform of oracle.xceff."""
    from . import xceff
    gga = GGA[xc_id]
    nvl = {"rho": 1, "sigma": 4, "tau": 5}[den_list]
    raw = eval_raw(xc_id, rho, nspin, 2)
    rho = np.asarray(rho)
    ng = raw.shape[0]
    exc = raw[:, 0].copy()
    if nspin == 1:
        vxc = np.zeros((ng, nvl))
        fxc = np.zeros((ng, nvl, nvl))
        if gga:
            vxc[:, :4] = xceff.transform(rho[:, :4], raw, "sigma", 1)
            fxc[:, :4, :4] = xceff.transform(rho[:, :4], raw, "sigma", 2)
        else:
            vxc[:, 0] = raw[:, 1]
            fxc[:, 0, 0] = raw[:, 2]
        return exc, vxc, fxc
    vxc = np.zeros((ng, nvl, 2))
    fxc = np.zeros((ng, nvl, 2, nvl, 2))
    den = "sigma" if gga else "rho"
    xcf = raw[:, xceff.libxc_to_xcfun_map(den, 2)]
    nv = 4 if gga else 1
    vxc[:, :nv] = xceff.transform_pol(rho[:, :nv], xcf, den, 1)
    fxc[:, :nv, :, :nv] = xceff.transform_pol(rho[:, :nv], xcf, den, 2)
    return exc, vxc, fxc
