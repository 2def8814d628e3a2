"""Pin the harmonic-analysis oracle: the reference's hard-coded NH3 fixtures (tests/test_vib.rs)
and Hessians built from known normal modes."""
import numpy as np
import pytest

from oracle import vib as ov

# tests/test_vib.rs:16-30: geometry [3, natm] (bohr), rows x, y, z; masses (u)
NH3_GEOM = np.array([[0.0, 1.88972612, 0.56691784, 0.18897261],
                     [0.0, 0.18897261, 2.07869874, 0.18897261],
                     [0.0, 0.37794522, 0.37794522, 2.26767135]])
NH3_MASS = np.array([14.007, 1.008, 1.008, 1.008])


def centred(geom, mass):
    return geom - (geom * mass).sum(1, keepdims=True) / mass.sum()


def test_nh3_rotation_constants_golden():
    """tests/test_vib.rs:100-120."""
    gc = centred(NH3_GEOM, NH3_MASS)
    ghz = ov.rotation_const(NH3_MASS, gc, "GHz")
    cm = ov.rotation_const(NH3_MASS, gc, "wavenumber")
    assert np.allclose(ghz, [285.07591188, 205.77561468, 181.01823364], atol=1e-5, rtol=0)
    assert np.allclose(cm, [9.50910889, 6.86393565, 6.03811833], atol=1e-7, rtol=0)
    assert ov.get_rotor_type(ghz) == "REGULAR"


def test_rotor_classification_golden():
    """tests/test_vib.rs:123-135."""
    assert ov.get_rotor_type([1e9, 1e9, 1e9]) == "ATOM"
    assert ov.get_rotor_type([5e8, 2.0, 2.0]) == "LINEAR"
    assert ov.get_rotor_type([300.0, 200.0, 180.0]) == "REGULAR"


def test_tr_space_shapes_golden():
    """tests/test_vib.rs:56-80 (geometries given row-major [3, natm] there)."""
    g1 = np.zeros((3, 1))
    assert ov.get_tr_space(np.ones(1), g1).shape == (3, 3)
    g2 = np.array([1.0, 0.0, 0.0, -1.0, 0.0, 0.0]).reshape(3, 2)
    assert ov.get_tr_space(np.ones(2), g2).shape == (6, 5)
    g4 = np.array([1.0, 1.0, 0.0, -1.0, 1.0, 0.0, -1.0, -1.0, 0.0, 1.0, -1.0, 0.0]).reshape(3, 4)
    tr = ov.get_tr_space(np.ones(4), g4)
    assert tr.shape == (12, 6)
    assert np.allclose(tr.T @ tr, np.eye(6), atol=1e-12, rtol=1e-12)
    tr = ov.get_tr_space(NH3_MASS, NH3_GEOM)
    assert tr.shape == (12, 6) and np.allclose(tr.T @ tr, np.eye(6), atol=1e-12, rtol=1e-12)


def test_rotation_constants_invariant_under_rotation():
    rng = np.random.default_rng(1)
    R, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    gc = centred(NH3_GEOM, NH3_MASS)
    assert np.allclose(ov.rotation_const(NH3_MASS, R @ gc), ov.rotation_const(NH3_MASS, gc), rtol=1e-12)


def closed_form(mass, lam, Q, res):
    """Frequencies, reduced masses, force constants and temperatures implied by the modes."""
    vib = res["vib"]
    order = np.argsort(lam)
    lam, Q = np.asarray(lam)[order], Q[:, order]
    om = np.sqrt(np.abs(lam)) * ov.uconv_cm1()
    w = Q / np.repeat(np.sqrt(mass), 3)[:, None]
    mu = 1.0 / (w ** 2).sum(0)
    return om, mu, vib


@pytest.mark.parametrize("seed", [0, 1])
def test_frequencies_of_constructed_hessian(seed):
    lam = np.array([0.02, 0.05, 0.11, 0.23, 0.31, 0.47]) * (1 + 0.1 * seed)
    hess, Q = ov.hessian_from_modes(NH3_MASS, NH3_GEOM, lam, seed=seed)
    r = ov.harmonic_analysis(hess, NH3_GEOM, NH3_MASS)
    assert r["nrt"] == 6 and r["trv"].count("TR") == 6 and r["trv"].count("V") == 6
    vi = [i for i, t in enumerate(r["trv"]) if t == "V"]
    assert np.allclose(r["omega"][vi], np.sqrt(lam) * ov.uconv_cm1(), rtol=1e-10)
    assert np.all(r["omega"][[i for i in range(12) if i not in vi]] < 1e-3)
    # reduced masses from the known modes (non-degenerate: unique up to sign)
    w = Q / np.repeat(np.sqrt(NH3_MASS), 3)[:, None]
    assert np.allclose(r["mu"][vi], 1.0 / (w ** 2).sum(0), rtol=1e-8)
    uconv_mdyne_a = 0.1 * (2 * np.pi * ov.C) ** 2 / ov.NA
    assert np.allclose(r["k"][vi], r["mu"][vi] * r["omega"][vi] ** 2 * uconv_mdyne_a, rtol=1e-12)
    assert np.allclose(r["theta_vib"][vi], r["omega"][vi] * 100 * ov.H * ov.C / ov.KB, rtol=1e-12)
    # the frequency of lam is independent of the masses' mode shapes: sqrt(lam) in a.u.
    assert np.all(np.abs(r["q"].T @ r["q"] - np.eye(12)) < 1e-10)


def test_projection_removes_tr_contamination():
    lam = np.array([0.03, 0.07, 0.13, 0.2, 0.29, 0.41])
    hess, _ = ov.hessian_from_modes(NH3_MASS, NH3_GEOM, lam, seed=3)
    tr = ov.get_tr_space(NH3_MASS, NH3_GEOM)
    rng = np.random.default_rng(4)
    D = rng.standard_normal((6, 6))
    sm = np.repeat(np.sqrt(NH3_MASS), 3)
    contam = (tr @ (D + D.T) @ tr.T) * sm[:, None] * sm[None, :]
    a = ov.harmonic_analysis(hess, NH3_GEOM, NH3_MASS)
    b = ov.harmonic_analysis(hess + contam, NH3_GEOM, NH3_MASS)
    vi = [i for i, t in enumerate(a["trv"]) if t == "V"]
    assert np.allclose(a["omega"][vi], b["omega"][vi], rtol=1e-10)


def test_linear_molecule_and_imaginary_mode():
    geom = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [-2.2, 0.0, 2.2]])  # O=C=O along z
    mass = np.array([15.999, 12.011, 15.999])
    lam = np.array([-0.01, 0.04, 0.09, 0.16])
    hess, _ = ov.hessian_from_modes(mass, geom, lam, seed=5, linear=True)
    r = ov.harmonic_analysis(hess, geom, mass)
    assert r["rotor"] == "LINEAR" and r["nrt"] == 5
    assert r["trv"].count("TR") == 5 and r["trv"].count("V") == 4
    assert r["imag"][0] and np.isclose(r["omega"][0], np.sqrt(0.01) * ov.uconv_cm1(), rtol=1e-10)
    assert r["k"][0] == 0.0 and r["theta_vib"][0] == 0.0
