"""Library harmonic analysis (host C++ behind the C ABI) against the oracle.  Runs on the CPU:
the hh_vib_* entry points make no GPU calls."""
import numpy as np
import pytest

from oracle import vib as ov
from tests.test_oracle_vib import NH3_GEOM, NH3_MASS, centred

vib = pytest.importorskip("hesship.vib")


def test_library_golden_rotation_constants():
    gc = centred(NH3_GEOM, NH3_MASS)
    ghz = vib.rotation_const(NH3_MASS, gc, "GHz")
    assert np.allclose(ghz, [285.07591188, 205.77561468, 181.01823364], atol=1e-5, rtol=0)
    assert np.allclose(vib.rotation_const(NH3_MASS, gc, "wavenumber"), [9.50910889, 6.86393565, 6.03811833],
                       atol=1e-7, rtol=0)
    assert vib.get_rotor_type(ghz) == "REGULAR"
    assert vib.get_rotor_type([1e9, 1e9, 1e9]) == "ATOM"
    assert vib.get_rotor_type([5e8, 2.0, 2.0]) == "LINEAR"


def test_library_tr_space():
    for mass, geom, ncol in [(np.ones(1), np.zeros((3, 1)), 3),
                             (np.ones(2), np.array([1.0, 0, 0, -1, 0, 0]).reshape(3, 2), 5),
                             (NH3_MASS, NH3_GEOM, 6)]:
        tr = vib.get_tr_space(mass, geom)
        assert tr.shape == (3 * len(mass), ncol)
        assert np.allclose(tr.T @ tr, np.eye(ncol), atol=1e-12)
        ref = ov.get_tr_space(mass, geom)
        assert np.allclose(tr @ tr.T, ref @ ref.T, atol=1e-12)  # same subspace


CASES = {
    "nh3": (NH3_MASS, NH3_GEOM, np.array([0.02, 0.05, 0.11, 0.23, 0.31, 0.47]), False),
    "nh3_imag": (NH3_MASS, NH3_GEOM, np.array([-0.015, 0.05, 0.11, 0.23, 0.31, 0.47]), False),
    "co2": (np.array([15.999, 12.011, 15.999]), np.array([[0.0, 0, 0], [0, 0, 0], [-2.2, 0, 2.2]]),
            np.array([0.04, 0.09, 0.16, 0.3]), True),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("proj", [(True, True), (True, False), (False, False)])
def test_library_matches_oracle(case, proj):
    mass, geom, lam, linear = CASES[case]
    hess, _ = ov.hessian_from_modes(mass, geom, lam, seed=7, linear=linear)
    rng = np.random.default_rng(8)
    hess = hess + 1e-4 * (lambda a: a + a.T)(rng.standard_normal(hess.shape))  # generic, TR not exact
    got = vib.harmonic_analysis(hess, geom, mass, *proj)
    ref = ov.harmonic_analysis(hess, geom, mass, *proj)
    # projected-out modes sit at the square root of rounding-level eigenvalues (~1e-5 cm^-1);
    # the reference accepts them below 1 cm^-1 (tests/test_vib.rs:157-160)
    assert np.allclose(got.omega, ref["omega"], rtol=1e-9, atol=1e-3)
    assert got.trv == ref["trv"]
    assert np.array_equal(got.degeneracy, ref["degeneracy"])
    v = np.array([t == "V" for t in ref["trv"]])
    assert np.array_equal(got.imag[v], ref["imag"][v])
    for f in ("mu", "k", "qtp0", "xtp0", "dq0", "theta_vib"):
        assert np.allclose(getattr(got, f)[v], ref[f][v], rtol=1e-8, atol=1e-12), f
    # normal modes of the (non-degenerate) vibrations agree column by column after the phase rule
    assert np.allclose(got.q[:, v], ref["q"][:, v], atol=1e-8)
    assert np.allclose(got.x[:, v], ref["x"][:, v], atol=1e-8)
    assert np.allclose(got.q.T @ got.q, np.eye(got.ndof), atol=1e-10)


def test_library_rejects_bad_input():
    from hesship._lib import HesshipError
    with pytest.raises(HesshipError):
        vib.harmonic_analysis(np.eye(12), NH3_GEOM, -NH3_MASS)
