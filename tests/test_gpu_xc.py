"""GPU parity: exchange-correlation functionals (hh_xc_eval, hh_xc_eval_eff) against the oracle.

Densities span 1e-5 .. 1e2 with reduced gradients from small to large, as on a molecular grid;
for the meta-GGAs tau spans tau_W .. 6 tau_W plus a uniform-gas part, with a tenth of the points
below tau_W (where TPSS takes tau = tau_W).
Tolerance (FP64): |gpu - oracle| <= 1e-9 |oracle| + 1e-11 * (largest magnitude among the same
point's values of the same derivative order).  The absolute part covers round-off in
cancellation-prone places (1 - zeta near 0, the LYP brace) whose conditioning is ~1e-10; every
value is otherwise held to 1e-9 relative.  Where the problem itself is ill-conditioned the
tolerance also holds 100 times the oracle's own change under relative perturbations of the
inputs of a few units in the last place (three draws): the TPSS correlation at strong spin
polarisation (second derivatives that are differences of terms 1e3 times larger), and in
one-orbital regions (tau <= tau_W) with both spin channels on eps_PBE, where it vanishes in exact
arithmetic (both implementations then give zero, or rounding below 1e-16: the absolute floor).
"""

import numpy as np
import pytest
import torch

from oracle import xcfunc as X
from tests.test_gpu_cphf import T

pytestmark = pytest.mark.gpu

IDS = sorted(X.FAMILY)
NAMES = {v: k for k, v in X.IDS.items()}
DEN = ("rho", "sigma", "tau")


def dens(rng, ng, lo=-5.0, hi=2.0, fam=1):
    r = np.empty((ng, 5 if fam == 2 else 4))
    r[:, 0] = 10 ** rng.uniform(lo, hi, ng)
    r[:, 1:4] = rng.normal(size=(ng, 3)) * r[:, :1] * rng.uniform(0.05, 3.0, (ng, 1))
    if fam == 2:
        tw = (r[:, 1:4] ** 2).sum(1) / (8 * r[:, 0])
        f = rng.uniform(1.02, 6.0, ng)
        f[rng.random(ng) < 0.1] = 0.5
        r[:, 4] = f * tw + np.where(f > 1, rng.uniform(0.0, 0.5, ng), 0.0) * r[:, 0] ** (5 / 3)
    return r


def case(nspin, ng=4000, seed=0, fam=1):
    rng = np.random.default_rng(seed)
    if nspin == 1:
        return dens(rng, ng, fam=fam)
    return np.stack([dens(rng, ng, fam=fam), dens(rng, ng, fam=fam)], -1)


def conditioning(fn, rho, seed=0):
    """max over three draws of |fn(rho (1 + 4e-16 u)) - fn(rho)|, u standard normal, for fn
    returning an array or a tuple of arrays."""
    rng = np.random.default_rng(seed)
    base = fn(rho)
    single = not isinstance(base, tuple)
    base = (base,) if single else base
    out = [np.zeros_like(b) for b in base]
    for _ in range(3):
        pert = fn(rho * (1 + 4e-16 * rng.standard_normal(rho.shape)))
        pert = (pert,) if single else pert
        out = [np.maximum(o, np.abs(p - b)) for o, p, b in zip(out, pert, base)]
    return out[0] if single else tuple(out)


def close(got, ref, blocks, cond=None):
    """blocks: column ranges of one derivative order; cond: the oracle's conditioning estimate
    (same shape as ref); returns the worst violation ratio."""
    worst = 0.0
    for lo, hi in blocks:
        g, r = got[:, lo:hi], ref[:, lo:hi]
        sc = np.maximum(np.abs(r).max(1, keepdims=True), 1e-5)   # absolute floor 1e-16
        tol = 1e-9 * np.abs(r) + 1e-11 * sc
        if cond is not None:
            tol = tol + 100 * cond[:, lo:hi]
        worst = max(worst, float((np.abs(g - r) / np.maximum(tol, 1e-300)).max()))
    return worst


def raw_blocks(xc_id, nspin, deriv):
    fam = X.FAMILY[xc_id]
    n = fam + 1 if nspin == 1 else (2, 5, 7)[fam]
    b = [(0, 1)]
    if deriv >= 1:
        b.append((1, 1 + n))
    if deriv >= 2:
        b.append((1 + n, 1 + n + n * (n + 1) // 2))
    return b


@pytest.mark.parametrize("nspin", [1, 2])
@pytest.mark.parametrize("xc_id", IDS)
def test_raw_against_oracle(dev, xc_id, nspin):
    import hesship
    rho = case(nspin, seed=xc_id + 10 * nspin, fam=X.FAMILY[xc_id])
    f = hesship.LibXCFunctional.from_identifier(NAMES[xc_id], "unpolarized" if nspin == 1 else "polarized")
    for deriv in (0, 1, 2):
        got = hesship.libxc_eval_inner(f, T(rho, dev), deriv).cpu().numpy()
        ref = X.eval_raw(xc_id, rho, nspin, deriv)
        assert got.shape == ref.shape == (rho.shape[0], X.ncol(xc_id, nspin, deriv))
        cond = conditioning(lambda q: X.eval_raw(xc_id, q, nspin, deriv), rho)
        assert close(got, ref, raw_blocks(xc_id, nspin, deriv), cond) <= 1.0, (NAMES[xc_id], nspin, deriv)


def _eff_close(got, ref, ng, cond=None):
    worst = 0.0
    for k, (g, r) in enumerate(zip(got, ref)):
        c = None if cond is None else cond[k].reshape(ng, -1)
        worst = max(worst, close(g.reshape(ng, -1), r.reshape(ng, -1), [(0, r.reshape(ng, -1).shape[1])], c))
    return worst


@pytest.mark.parametrize("nspin", [1, 2])
@pytest.mark.parametrize("xc_id", IDS)
def test_eff_against_oracle(dev, xc_id, nspin):
    import hesship
    rho = case(nspin, ng=3000, seed=100 + xc_id + nspin, fam=X.FAMILY[xc_id])
    f = hesship.LibXCFunctional.from_identifier(NAMES[xc_id], "unpolarized" if nspin == 1 else "polarized")
    exc, vxc, fxc = hesship.libxc_eval_eff(f, T(rho, dev), 2)
    den = DEN[X.FAMILY[xc_id]]
    ref = X.eval_eff(xc_id, rho, nspin, den)
    got = [t.cpu().numpy() for t in (exc, vxc, fxc)]
    for g, r in zip(got, ref):
        assert g.shape == r.shape
    cond = conditioning(lambda q: X.eval_eff(xc_id, q, nspin, den), rho)
    assert _eff_close(got, ref, rho.shape[0], cond) <= 1.0, (NAMES[xc_id], nspin)


@pytest.mark.parametrize("nspin", [1, 2])
def test_eff_equals_raw_through_device_transform(dev, nspin):
    """The fused chain rule equals raw LibXC columns -> XCFun order -> hh_xceff_transform*."""
    import hesship
    rho = case(nspin, ng=2500, seed=7 + nspin)
    f = hesship.LibXCFunctional.from_identifier("HYB_GGA_XC_B3LYP", "unpolarized" if nspin == 1 else "polarized")
    d_rho = T(rho, dev)
    raw = hesship.libxc_eval_inner(f, d_rho, 2)
    _, vxc, fxc = hesship.libxc_eval_eff(f, d_rho, 2)
    if nspin == 1:
        v2 = hesship.transform_xc(d_rho, raw, "sigma", 1)
        f2 = hesship.transform_xc(d_rho, raw, "sigma", 2)
    else:
        m = torch.as_tensor(hesship.libxc_to_xcfun_mapping("sigma", 2), device=dev)
        xcf = raw[:, m]
        v2 = hesship.transform_xc_polarized(d_rho, xcf, "sigma", 1)
        f2 = hesship.transform_xc_polarized(d_rho, xcf, "sigma", 2)
    ng = rho.shape[0]
    assert _eff_close([vxc.cpu().numpy(), fxc.cpu().numpy()], [v2.cpu().numpy(), f2.cpu().numpy()], ng) <= 1.0


@pytest.mark.parametrize("nspin", [1, 2])
def test_functional_lists(dev, nspin):
    """get_rho_vxc_fxc's loop: SVWN (LDA_X + LDA_C_VWN) into RHO arrays; B3LYP as one functional
    equals its four weighted parts; an LDA term inside a SIGMA list fills the leading slice."""
    import hesship
    spin = "unpolarized" if nspin == 1 else "polarized"
    F = hesship.LibXCFunctional.from_identifier
    rho = case(nspin, ng=2000, seed=31 + nspin)
    d_rho = T(rho, dev)
    ng = rho.shape[0]
    # SVWN
    exc, vxc, fxc = hesship.eval_xc_eff_list([(1.0, F("LDA_X", spin)), (1.0, F("LDA_C_VWN", spin))], d_rho)
    ref = [a + b for a, b in zip(X.eval_eff(1, rho, nspin, "rho"), X.eval_eff(7, rho, nspin, "rho"))]
    assert _eff_close([t.cpu().numpy() for t in (exc, vxc, fxc)], ref, ng) <= 1.0
    # B3LYP by parts
    whole = hesship.eval_xc_eff_list([(1.0, F("HYB_GGA_XC_B3LYP", spin))], d_rho)
    parts = hesship.eval_xc_eff_list([(0.08, F("LDA_X", spin)), (0.72, F("GGA_X_B88", spin)),
                                      (0.19, F("LDA_C_VWN_RPA", spin)), (0.81, F("GGA_C_LYP", spin))], d_rho)
    assert _eff_close([t.cpu().numpy() for t in whole], [t.cpu().numpy() for t in parts], ng) <= 1.0
    assert F("HYB_GGA_XC_B3LYP", spin).hyb_exx_coef() == 0.2
    # leading slice
    _, v_l, f_l = hesship.eval_xc_eff_list([(1.0, F("LDA_C_VWN", spin)), (0.0, F("GGA_X_B88", spin))], d_rho)
    ref_l = X.eval_eff(7, rho, nspin, "sigma")
    assert _eff_close([v_l.cpu().numpy(), f_l.cpu().numpy()], list(ref_l[1:]), ng) <= 1.0


def test_thresholds_and_zero_gradient(dev):
    import hesship
    F = hesship.LibXCFunctional.from_identifier
    # unpolarised: zero / sub-threshold density gives zeros; zero gradient gives B88 = Slater and
    # vsigma = -beta/2 (rho/2)^(-4/3) (the gradient-expansion coefficient)
    r = np.zeros((5, 4))
    r[:, 0] = [0.0, 5e-16, 0.3, 1.0, 4.0]
    got = hesship.libxc_eval_inner(F("GGA_X_B88"), T(r, dev), 2).cpu().numpy()
    assert np.all(got[:2] == 0.0)
    rho = r[2:, 0]
    np.testing.assert_allclose(got[2:, 0], -0.75 * (3 / np.pi) ** (1 / 3) * rho ** (1 / 3), rtol=1e-13)
    np.testing.assert_allclose(got[2:, 2], -0.0042 / 2 * (rho / 2) ** (-4 / 3), rtol=1e-13)
    # accumulate leaves sub-threshold points untouched
    exc, vxc, fxc = hesship.eval_xc_eff_list([(1.0, F("HYB_GGA_XC_B3LYP"))], T(r, dev))
    assert torch.all(exc[:2] == 0) and torch.all(vxc[:2] == 0) and torch.all(fxc[:2] == 0)
    # polarised with one empty spin: exchange of the occupied spin only, LYP vanishes
    rp = np.zeros((3, 4, 2))
    rp[:, 0, 0] = [0.2, 1.0, 3.0]
    rp[:, 1:, 0] = [[0.1, 0.0, 0.05], [0.3, -0.2, 0.1], [1.0, 0.5, -0.4]]
    g_x = hesship.libxc_eval_inner(F("GGA_X_B88", "polarized"), T(rp, dev), 0).cpu().numpy()[:, 0]
    g_u = hesship.libxc_eval_inner(F("GGA_X_B88"), T(2 * rp[:, :, 0], dev), 0).cpu().numpy()[:, 0]
    np.testing.assert_allclose(g_x * rp[:, 0, 0], 0.5 * g_u * 2 * rp[:, 0, 0], rtol=1e-13)
    g_c = hesship.libxc_eval_inner(F("GGA_C_LYP", "polarized"), T(rp, dev), 0).cpu().numpy()[:, 0]
    assert np.abs(g_c).max() < 1e-12


def test_errors(dev):
    import hesship
    from hesship._lib import HesshipError, lib
    with pytest.raises(ValueError):
        hesship.LibXCFunctional.from_identifier("MGGA_X_SCAN")
    with pytest.raises(HesshipError):
        hesship.LibXCFunctional(999)
    r = T(np.ones((4, 4)), dev)
    # GGA functional with RHO-only outputs is rejected
    f = hesship.LibXCFunctional.from_identifier("GGA_X_B88")
    with pytest.raises(HesshipError):
        from hesship.xc import _accumulate_eff
        out = torch.zeros(4, dtype=torch.float64, device=dev)
        _accumulate_eff(f, r, "rho", 1.0, out, out, out, False)
    # a meta-GGA needs TAU outputs
    with pytest.raises(HesshipError):
        out = torch.zeros(4, dtype=torch.float64, device=dev)
        _accumulate_eff(hesship.LibXCFunctional.from_identifier("MGGA_C_TPSS"), T(np.ones((4, 5)), dev), "sigma",
                        1.0, out, out, out, False)
    # empty input is a no-op
    assert lib().hh_xc_eval(1, 1, 2, None, 0, 0, 0, None, 0, None) == 0


@pytest.mark.parametrize("nspin", [1, 2])
def test_full_grid_size(dev, nspin):
    """B3LYP at the benchmark grid size (543600 points): finite everywhere and a 2000-point
    sample equal to the oracle."""
    import hesship
    ng = 543600
    rng = np.random.default_rng(77 + nspin)
    rho = case(nspin, ng=ng, seed=77 + nspin)
    f = hesship.LibXCFunctional.from_identifier("HYB_GGA_XC_B3LYP", "unpolarized" if nspin == 1 else "polarized")
    exc, vxc, fxc = hesship.libxc_eval_eff(f, T(rho, dev), 2)
    torch.cuda.synchronize()
    assert torch.isfinite(fxc).all() and torch.isfinite(vxc).all() and torch.isfinite(exc).all()
    idx = np.sort(rng.choice(ng, 2000, replace=False))
    ref = X.eval_eff(402, rho[idx], nspin, "sigma")
    got = [t.cpu().numpy()[idx] for t in (exc, vxc, fxc)]
    assert _eff_close(got, ref, len(idx)) <= 1.0


@pytest.mark.parametrize("nspin", [1, 2])
def test_tpss0_list_and_mixed_families(dev, nspin):
    """TPSS0 as one functional equals 0.75 TPSS_X + TPSS_C as a list; a list mixing B3LYP (GGA)
    and TPSS_C (meta-GGA) fills TAU arrays, the GGA term in the leading (rho, grad) slice."""
    import hesship
    F = hesship.LibXCFunctional.from_identifier
    spin = "unpolarized" if nspin == 1 else "polarized"
    rho = case(nspin, ng=2000, seed=31 + nspin, fam=2)
    d_rho = T(rho, dev)
    f0 = F("HYB_MGGA_XC_TPSS0", spin)
    assert f0.hyb_exx_coef() == 0.25 and f0.den_type == "tau"
    whole = hesship.eval_xc_eff_list([(1.0, f0)], d_rho)
    parts = hesship.eval_xc_eff_list([(0.75, F("MGGA_X_TPSS", spin)), (1.0, F("MGGA_C_TPSS", spin))], d_rho)
    for a, b in zip(whole, parts):
        torch.testing.assert_close(a, b, rtol=1e-12, atol=1e-12 * float(b.abs().max()))
    exc, vxc, fxc = hesship.eval_xc_eff_list([(1.0, F("HYB_GGA_XC_B3LYP", spin)), (1.0, F("MGGA_C_TPSS", spin))],
                                             d_rho)
    def both(q):
        r1 = X.eval_eff(402, q[:, :4], nspin, "tau")
        r2 = X.eval_eff(231, q, nspin, "tau")
        return tuple(a + b for a, b in zip(r1, r2))
    ref = both(rho)
    got = [t.cpu().numpy() for t in (exc, vxc, fxc)]
    assert _eff_close(got, ref, rho.shape[0], conditioning(both, rho)) <= 1.0


@pytest.mark.parametrize("nspin", [1, 2])
def test_tpss0_full_grid_size(dev, nspin):
    """TPSS0 at the benchmark grid size: finite everywhere, a 1500-point sample equal to the oracle."""
    import hesship
    ng = 543600
    rng = np.random.default_rng(91 + nspin)
    rho = case(nspin, ng=ng, seed=91 + nspin, fam=2)
    f = hesship.LibXCFunctional.from_identifier("HYB_MGGA_XC_TPSS0", "unpolarized" if nspin == 1 else "polarized")
    exc, vxc, fxc = hesship.libxc_eval_eff(f, T(rho, dev), 2)
    torch.cuda.synchronize()
    assert torch.isfinite(fxc).all() and torch.isfinite(vxc).all() and torch.isfinite(exc).all()
    idx = np.sort(rng.choice(ng, 1500, replace=False))
    ref = X.eval_eff(396, rho[idx], nspin, "tau")
    got = [t.cpu().numpy()[idx] for t in (exc, vxc, fxc)]
    cond = conditioning(lambda q: X.eval_eff(396, q, nspin, "tau"), rho[idx])
    assert _eff_close(got, ref, len(idx), cond) <= 1.0
