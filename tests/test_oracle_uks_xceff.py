"""Pin the spin-polarised xceff transform, the LibXC -> XCFun column map and the UKS grid
response of the oracle (CPU)."""
import numpy as np
import pytest
import torch

from oracle import numint as oni
from oracle import xceff as oxc

# src/xceff/xc_deriv.rs:17-40 (libxc_to_xcfun_mapping_parts), orders 0-2
REF_MAP = {
    "sigma": [0] + [1, 2, 3, 4, 5] + [6, 7, 9, 10, 11, 8, 12, 13, 14, 15, 16, 17, 18, 19, 20],
    "tau": [0] + [1, 2, 3, 4, 5, 6, 7] + [8, 9, 11, 12, 13, 17, 18, 10, 14, 15, 16, 19, 20, 21, 22, 23, 27, 28, 24, 25,
                                        29, 30, 26, 31, 32, 33, 34, 35],
}


@pytest.mark.parametrize("den", ["sigma", "tau"])
def test_libxc_map_matches_reference_table(den):
    assert oxc.libxc_to_xcfun_map(den, 2) == REF_MAP[den]


def energy_pol(ra, rb, saa, sab, sbb, ta=None, tb=None):
    """A smooth analytic spin-polarised energy density (not a physical functional)."""
    e = -0.6 * (ra ** (4.0 / 3.0) + rb ** (4.0 / 3.0)) + 0.03 * (saa + sbb) / (1.0 + ra + rb)
    e = e + 0.02 * sab * torch.exp(-0.3 * (ra + rb)) + 0.01 * torch.log1p(saa * sbb) * ra * rb
    if ta is not None:
        e = e + 0.04 * ta * ra / (1.0 + ta) + 0.03 * tb * rb / (1.0 + tb) + 0.002 * (ta * sbb + tb * saa)
    return e


@pytest.mark.parametrize("den", ["sigma", "tau"])
def test_polarised_chain_rule_against_autograd(den):
    rng = np.random.default_rng(5)
    ng = 6
    nv, xlen = oxc.NVAR_POL[den], oxc.XLEN_POL[den]
    rho = np.zeros((ng, nv, 2))
    rho[:, 0, :] = rng.uniform(0.2, 1.4, (ng, 2))
    rho[:, 1:4, :] = rng.standard_normal((ng, 3, 2)) * 0.4
    if den == "tau":
        rho[:, 4, :] = rng.uniform(0.1, 1.0, (ng, 2))
    xc = np.zeros((ng, 1 + xlen + xlen * (xlen + 1) // 2))
    vref = np.zeros((ng, nv, 2))
    fref = np.zeros((ng, nv, 2, nv, 2))
    for g in range(ng):
        ga, gb = rho[g, 1:4, 0], rho[g, 1:4, 1]
        u = [rho[g, 0, 0], rho[g, 0, 1], ga @ ga, ga @ gb, gb @ gb]
        if den == "tau":
            u += [rho[g, 4, 0], rho[g, 4, 1]]
        uv = torch.tensor(u, dtype=torch.float64, requires_grad=True)
        f = lambda q: energy_pol(*q)
        d1 = torch.autograd.functional.jacobian(f, uv).numpy()
        d2 = torch.autograd.functional.hessian(f, uv).numpy()
        xc[g, 0] = f(uv).item()
        xc[g, 1:1 + xlen] = d1
        for k, (p, q) in enumerate(oxc.uniq_pairs(xlen)):
            xc[g, 1 + xlen + k] = d2[p, q]
        # direct derivatives in (rho_s, grad_s, tau_s), flattened (x, s) with x fastest
        r = torch.tensor(rho[g].T.reshape(-1), dtype=torch.float64, requires_grad=True)  # [(x, s)] s-major

        def h(v):
            a, b = v[:nv], v[nv:]
            args = [a[0], b[0], a[1:4] @ a[1:4], a[1:4] @ b[1:4], b[1:4] @ b[1:4]]
            if den == "tau":
                args += [a[4], b[4]]
            return energy_pol(*args)
        vj = torch.autograd.functional.jacobian(h, r).numpy().reshape(2, nv).T
        vh = torch.autograd.functional.hessian(h, r).numpy().reshape(2, nv, 2, nv).transpose(1, 0, 3, 2)
        vref[g], fref[g] = vj, vh
    np.testing.assert_allclose(oxc.transform_pol(rho, xc, den, 1), vref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oxc.transform_pol(rho, xc, den, 2), fref, rtol=1e-11, atol=1e-12)


def test_polarised_lda_is_the_raw_block():
    rng = np.random.default_rng(1)
    xc = rng.standard_normal((5, 1 + 2 + 3))
    out = oxc.transform_pol(np.zeros((5, 1, 2)), xc, "rho", 2)
    # product_uniq_indices(2, 2) = [0, 1, 1, 2] over (aa, ab, bb)
    np.testing.assert_array_equal(out[:, 0, :, 0, :].reshape(5, 4), xc[:, 3:][:, [0, 1, 1, 2]])


def uks_case(den, ng=13, nao=6, nocc=(2, 3), nset=2, seed=0):
    rng = np.random.default_rng(seed)
    ao = rng.standard_normal((ng, nao, 4))
    w = rng.uniform(0.1, 1.0, ng)
    mocc = [rng.standard_normal((nao, n)) for n in nocc]
    kets = [rng.standard_normal((nao, n, nset)) for n in nocc]
    nv = oni.NVAR[den]
    f = rng.standard_normal((ng, nv * 2, nv * 2))
    f = (f + f.transpose(0, 2, 1)).reshape(ng, 2, nv, 2, nv).transpose(0, 2, 1, 4, 3)  # [g, x, s, y, t]
    return ao, w, mocc, kets, f


@pytest.mark.parametrize("den", ["rho", "sigma", "tau"])
def test_uks_bra_transformed_potential_equals_full_potential_times_bra(den):
    ao, w, mocc, kets, f = uks_case(den)
    rho1 = np.stack([oni.brute_rho1(ao, mocc[s], kets[s], den) for s in range(2)], axis=2)
    V = oni.brute_uks_potential(den, f, rho1, ao, w)
    out = oni.uks_fxc_pot_bra_trans(den, f, rho1, ao, w, mocc)
    for s in range(2):
        np.testing.assert_allclose(out[s], np.einsum("uva,vi->uia", V[s], mocc[s]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("den", ["sigma", "tau"])
def test_uks_response_couples_the_spins(den):
    ao, w, mocc, kets, f = uks_case(den, seed=2)
    r1 = oni.uks_response_bra(den, f, ao, w, kets, mocc)
    r2 = oni.uks_response_bra(den, f, ao, w, [kets[0], 0.0 * kets[1]], mocc)
    assert np.abs(r2[1]).max() > 1e-3  # alpha perturbation reaches the beta response
    # linearity in the pair of perturbations
    r3 = oni.uks_response_bra(den, f, ao, w, [2.0 * kets[0], -kets[1]], mocc)
    ra = oni.uks_response_bra(den, f, ao, w, [kets[0], 0 * kets[1]], mocc)
    rb = oni.uks_response_bra(den, f, ao, w, [0 * kets[0], kets[1]], mocc)
    for s in range(2):
        np.testing.assert_allclose(r3[s], 2.0 * ra[s] - rb[s], rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(r1[s], ra[s] + rb[s], rtol=1e-11, atol=1e-11)
