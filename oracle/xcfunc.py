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

IDS = {"LDA_X": 1, "LDA_C_VWN": 7, "LDA_C_VWN_RPA": 8, "GGA_X_B88": 106, "GGA_C_LYP": 131, "HYB_GGA_XC_B3LYP": 402,
       "MGGA_X_TPSS": 202, "MGGA_C_TPSS": 231, "HYB_MGGA_XC_TPSS0": 396}
# family: 0 LDA (rho), 1 GGA (rho, sigma), 2 meta-GGA (rho, sigma, tau)
FAMILY = {1: 0, 7: 0, 8: 0, 106: 1, 131: 1, 402: 1, 202: 2, 231: 2, 396: 2}
GGA = {k: v >= 1 for k, v in FAMILY.items()}
EXX = {402: 0.2, 396: 0.25}

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


# ---- TPSS (Tao, Perdew, Staroverov, Scuseria, Phys. Rev. Lett. 91, 146401 (2003)) ----------
TPSS_KAPPA, TPSS_B, TPSS_C, TPSS_E, TPSS_MU, TPSS_D = 0.804, 0.40, 1.59096, 1.537, 0.21951, 2.8
K2 = (3.0 * math.pi ** 2) ** (2.0 / 3.0)


def tpss_x_total(n, s, tau):
    """TPSS exchange energy per volume of a spin-unpolarised density n (|grad n|^2 = s, kinetic
    energy density tau): e_x^unif(n) F_x(p, z) with p = s / (4 (3 pi^2)^(2/3) n^(8/3)),
    z = tau_W / tau (tau_W = s / 8n; tau below tau_W taken as tau_W), alpha = (tau - tau_W) /
    tau_unif, q_b = 9/20 (alpha - 1) / sqrt(1 + b alpha (alpha - 1)) + 2p/3 and
    x = {[10/81 + c z^2/(1+z^2)^2] p + 146/2025 q_b^2 - 73/405 q_b sqrt(1/2 (3z/5)^2 + 1/2 p^2)
         + (10/81)^2 p^2 / kappa + 2 sqrt(e) 10/81 (3z/5)^2 + e mu p^3} / (1 + sqrt(e) p)^2,
    F_x = 1 + kappa - kappa / (1 + x / kappa).  The square root is taken as p sqrt(1/2 (3r/5)^2 + 1/2)
    with r = z / p = (3 pi^2)^(2/3) n^(5/3) / (2 tau), which has no kink at zero gradient."""
    exunif = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * n ** (4.0 / 3.0)
    p = s / (4.0 * K2 * n ** (8.0 / 3.0))
    tw = s / (8.0 * n)
    above = tau > tw
    taue = torch.where(above, tau, tw)
    # z = 1 and alpha = 0 exactly (constants) where tau is taken as tau_W
    z = torch.where(above, tw / taue, torch.ones_like(tw))
    alpha = torch.where(above, (taue - tw) / (0.3 * K2 * n ** (5.0 / 3.0)), torch.zeros_like(tw))
    qb = 0.45 * (alpha - 1.0) / torch.sqrt(1.0 + TPSS_B * alpha * (alpha - 1.0)) + 2.0 * p / 3.0
    r = K2 * n ** (5.0 / 3.0) / (2.0 * taue)
    se = math.sqrt(TPSS_E)
    x = ((10.0 / 81.0 + TPSS_C * z ** 2 / (1.0 + z ** 2) ** 2) * p + 146.0 / 2025.0 * qb ** 2
         - 73.0 / 405.0 * qb * p * torch.sqrt(0.5 * (0.6 * r) ** 2 + 0.5)
         + (10.0 / 81.0) ** 2 / TPSS_KAPPA * p ** 2 + 2.0 * se * 10.0 / 81.0 * (0.6 * z) ** 2
         + TPSS_E * TPSS_MU * p ** 3) / (1.0 + se * p) ** 2
    return exunif * (1.0 + TPSS_KAPPA - TPSS_KAPPA / (1.0 + x / TPSS_KAPPA))


def tpss_x_spin(r, s, t):
    """Exchange of one spin channel by spin scaling: 1/2 e_x(2 rho_s, 4 sigma_ss, 2 tau_s)."""
    return 0.5 * tpss_x_total(2.0 * r, 4.0 * s, 2.0 * t)


def _opz(t, p):
    return torch.where(t <= ZETA_THR, torch.full_like(t, ZETA_THR ** p), torch.clamp(t, min=ZETA_THR) ** p)


PWMOD = [(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294),
         (0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517),
         (0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671)]   # paramagnetic, ferromagnetic, -alpha_c
PW_FPP = 1.709920934161365617563962776245


def pw92_g(rs, A, a1, b1, b2, b3, b4):
    srs = torch.sqrt(rs)
    return -2.0 * A * (1.0 + a1 * rs) * torch.log1p(1.0 / (2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs)))


def pw92_eps(rs, z, opz=None, omz=None):
    """Perdew-Wang 1992 correlation per particle (the constants of LibXC's modified version);
    opz, omz: 1 + zeta and 1 - zeta when the caller has them without cancellation."""
    opz = 1.0 + z if opz is None else opz
    omz = 1.0 - z if omz is None else omz
    e0, e1 = pw92_g(rs, *PWMOD[0]), pw92_g(rs, *PWMOD[1])
    ac = -pw92_g(rs, *PWMOD[2])
    f = (_opz(opz, 4.0 / 3.0) + _opz(omz, 4.0 / 3.0) - 2.0) / (2.0 ** (4.0 / 3.0) - 2.0)
    z4 = z ** 4
    return e0 + ac * f * (1.0 - z4) / PW_FPP + (e1 - e0) * f * z4


PBE_BETA = 0.06672455060314922
PBE_GAMMA = (1.0 - math.log(2.0)) / math.pi ** 2


def pbe_c_eps(n, z, s, opz=None, omz=None):
    """PBE correlation per particle (Perdew, Burke, Ernzerhof 1996): PW92 + H(rs, zeta, t),
    H = gamma phi^3 ln[1 + beta/gamma t^2 (1 + A t^2) / (1 + A t^2 + A^2 t^4)],
    A = beta/gamma / (exp(-eps_c / (gamma phi^3)) - 1), t^2 = pi s / (16 phi^2 k_F n^2)."""
    opz = 1.0 + z if opz is None else opz
    omz = 1.0 - z if omz is None else omz
    rs = (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0) * n ** (-1.0 / 3.0)
    ec = pw92_eps(rs, z, opz, omz)
    phi = (_opz(opz, 2.0 / 3.0) + _opz(omz, 2.0 / 3.0)) / 2.0
    kf = (3.0 * math.pi ** 2 * n) ** (1.0 / 3.0)
    t2 = math.pi * s / (16.0 * phi ** 2 * kf * n ** 2)
    phi3 = phi ** 3
    A = PBE_BETA / PBE_GAMMA / torch.expm1(-ec / (PBE_GAMMA * phi3))
    y = A * t2
    return ec + PBE_GAMMA * phi3 * torch.log1p(PBE_BETA / PBE_GAMMA * t2 * (1.0 + y) / (1.0 + y + y * y))


def tpss_c(ra, rb, saa, sab, sbb, ta, tb):
    """TPSS correlation energy per volume: n eps_rev (1 + d eps_rev z^3), eps_rev = eps_PBE (1 + C z^2)
    - (1 + C) z^2 sum_s n_s / n max(eps_PBE(n_s, 0), eps_PBE), C(zeta, xi) = (0.53 + 0.87 zeta^2
    + 0.50 zeta^4 + 2.26 zeta^6) / {1 + xi^2 [(1+zeta)^(-4/3) + (1-zeta)^(-4/3)] / 2}^4,
    xi = |grad zeta| / (2 (3 pi^2 n)^(1/3)), z = tau_W / tau, d = 2.8."""
    n = ra + rb
    z = (ra - rb) / n
    opz, omz = 2.0 * ra / n, 2.0 * rb / n   # 1 +- zeta without cancellation
    s = saa + 2.0 * sab + sbb
    epbe = pbe_c_eps(n, z, s, opz, omz)
    one = torch.ones_like(ra)
    ea, eb = pbe_c_eps(ra, one, saa), pbe_c_eps(rb, one, sbb)
    eta = torch.where(ea > epbe, ea, epbe)   # max, the second argument on a tie
    etb = torch.where(eb > epbe, eb, epbe)
    tw = s / (8.0 * n)
    tt = ta + tb
    zt = torch.where(tt > tw, tw / tt, torch.ones_like(tw))
    gz2 = 4.0 / n ** 4 * (rb ** 2 * saa - 2.0 * ra * rb * sab + ra ** 2 * sbb)
    xi2 = gz2 / (4.0 * (3.0 * math.pi ** 2 * n) ** (2.0 / 3.0))
    c0 = 0.53 + 0.87 * z ** 2 + 0.50 * z ** 4 + 2.26 * z ** 6
    cz = c0 / (1.0 + xi2 * (_opz(opz, -4.0 / 3.0) + _opz(omz, -4.0 / 3.0)) / 2.0) ** 4
    # eps_PBE (1 + C z^2) - (1 + C) z^2 sum_s n_s/n eps_s, written with sum_s n_s/n = 1 so that
    # it is zero exactly (value and derivatives) where z = 1 and eps_s = eps_PBE for both spins
    erev = epbe * (1.0 - zt ** 2) + (1.0 + cz) * zt ** 2 * (ra / n * (epbe - eta) + rb / n * (epbe - etb))
    # there it vanishes identically in a neighbourhood; the explicit zero keeps autograd's
    # second-order residue of the cancelling terms (~1e-14) out of the derivatives
    zero = (tt <= tw) & ~(ea > epbe) & ~(eb > epbe)
    erev = torch.where(zero, torch.zeros_like(erev), erev)
    return n * erev * (1.0 + TPSS_D * erev * zt ** 3)


def energy_density(xc_id, ra, rb, saa, sab, sbb, on_a, on_b, ta=None, tb=None):
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

    def tpssx():
        return (torch.where(on_a, tpss_x_spin(ra, saa, ta), 0.0 * ra)
                + torch.where(on_b, tpss_x_spin(rb, sbb, tb), 0.0 * rb))

    if xc_id == 202:
        return tpssx()
    if xc_id == 231:
        return tpss_c(ra, rb, saa, sab, sbb, ta, tb)
    if xc_id == 396:
        return 0.75 * tpssx() + tpss_c(ra, rb, saa, sab, sbb, ta, tb)
    raise ValueError(f"unsupported functional id {xc_id}")


def _variables(rho, nspin, fam):
    """LibXC variables from rho [ng, nvar] (nspin 1) or [ng, nvar, 2] (nspin 2): rho(s), then
    sigma(s) (GGA, meta-GGA), then tau(s) (meta-GGA, component 4)."""
    rho = torch.as_tensor(np.asarray(rho), dtype=torch.float64)
    if nspin == 1:
        v = [rho[:, 0].clone()]
        if fam >= 1:
            v.append(torch.clamp((rho[:, 1:4] ** 2).sum(1), min=SIGMA_FLOOR))
        if fam >= 2:
            v.append(rho[:, 4].clone())
        rtot = rho[:, 0]
    else:
        v = [rho[:, 0, 0].clone(), rho[:, 0, 1].clone()]
        if fam >= 1:
            ga, gb = rho[:, 1:4, 0], rho[:, 1:4, 1]
            v += [torch.clamp((ga * ga).sum(1), min=SIGMA_FLOOR), (ga * gb).sum(1),
                  torch.clamp((gb * gb).sum(1), min=SIGMA_FLOOR)]
        if fam >= 2:
            v += [rho[:, 4, 0].clone(), rho[:, 4, 1].clone()]
        rtot = rho[:, 0, 0] + rho[:, 0, 1]
    return v, rtot


def _polarised_args(v, nspin, fam):
    """(ra, rb, saa, sab, sbb, on_a, on_b, ta, tb) of the polarised energy densities."""
    if nspin == 1:
        r = v[0]
        ra = rb = 0.5 * r
        s = 0.25 * v[1] if fam >= 1 else torch.zeros_like(r)
        t = 0.5 * v[2] if fam >= 2 else torch.zeros_like(r)
        on = torch.ones_like(r, dtype=torch.bool)
        return ra, rb, s, s, s, on, on, t, t
    on_a, on_b = v[0] > DENS_THR, v[1] > DENS_THR
    ra = torch.clamp(v[0], min=DENS_THR)
    rb = torch.clamp(v[1], min=DENS_THR)
    z = torch.zeros_like(ra)
    sg = (v[2], v[3], v[4]) if fam >= 1 else (z, z, z)
    tt = (v[5], v[6]) if fam >= 2 else (z, z)
    return (ra, rb) + sg + (on_a, on_b) + tt


def libxc_order_pairs(nvar, nspin, gga):
    """Second-derivative variable pairs in LibXC's column order (polarised: v2rho2, v2rhosigma,
    v2rhotau, v2sigma2, v2sigmatau, v2tau2 blocks, each spin-pair list in row-major order)."""
    if nspin == 1 or not gga:
        return [(i, j) for i in range(nvar) for j in range(i, nvar)]
    pairs = ([(0, 0), (0, 1), (1, 1)] + [(i, j) for i in range(2) for j in range(2, 5)])
    if nvar == 7:
        pairs += [(i, j) for i in range(2) for j in range(5, 7)]
    pairs += [(i, j) for i in range(2, 5) for j in range(i, 5)]
    if nvar == 7:
        pairs += [(i, j) for i in range(2, 5) for j in range(5, 7)] + [(5, 5), (5, 6), (6, 6)]
    return pairs


def eval_raw(xc_id, rho, nspin, deriv):
    """LibXC output [ng, ncol] (numpy) of functional xc_id at rho."""
    gga = GGA[xc_id]
    fam = FAMILY[xc_id]
    v, rtot = _variables(rho, nspin, fam)
    for t in v:
        t.requires_grad_(True)
    ra, rb, saa, sab, sbb, on_a, on_b, ta, tb = _polarised_args(v, nspin, fam)
    e = energy_density(xc_id, ra, rb, saa, sab, sbb, on_a, on_b, ta, tb)
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
    n = (1 + FAMILY[xc_id]) if nspin == 1 else (2, 5, 7)[FAMILY[xc_id]]
    return 1 + (n if deriv >= 1 else 0) + (n * (n + 1) // 2 if deriv >= 2 else 0)


def eval_eff(xc_id, rho, nspin, den_list):
    """libxc_eval_eff at order 2 embedded in the leading slice of the caller's arrays:
    (exc [ng], vxc [ng, nvl(, 2)], fxc [ng, nvl(, 2), nvl(, 2)]) with nvl that of den_list.
    Chain rule by the separately pinned transform of oracle.xceff."""
    from . import xceff
    gga = GGA[xc_id]
    fam = FAMILY[xc_id]
    nvl = {"rho": 1, "sigma": 4, "tau": 5}[den_list]
    raw = eval_raw(xc_id, rho, nspin, 2)
    rho = np.asarray(rho)
    ng = raw.shape[0]
    exc = raw[:, 0].copy()
    den = ("rho", "sigma", "tau")[fam]
    nv = (1, 4, 5)[fam]
    if nspin == 1:
        vxc = np.zeros((ng, nvl))
        fxc = np.zeros((ng, nvl, nvl))
        if gga:
            vxc[:, :nv] = xceff.transform(rho[:, :nv], raw, den, 1)
            fxc[:, :nv, :nv] = xceff.transform(rho[:, :nv], raw, den, 2)
        else:
            vxc[:, 0] = raw[:, 1]
            fxc[:, 0, 0] = raw[:, 2]
        return exc, vxc, fxc
    vxc = np.zeros((ng, nvl, 2))
    fxc = np.zeros((ng, nvl, 2, nvl, 2))
    xcf = raw[:, xceff.libxc_to_xcfun_map(den, 2)]
    vxc[:, :nv] = xceff.transform_pol(rho[:, :nv], xcf, den, 1)
    fxc[:, :nv, :, :nv] = xceff.transform_pol(rho[:, :nv], xcf, den, 2)
    return exc, vxc, fxc
