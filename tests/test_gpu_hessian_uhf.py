"""End to end on the device for open shells: density-fitted UHF / UKS (hh_scf_uhf) and the whole
analytic unrestricted Hessian (UHessSCF::make_hess, uscf.rs:514-540: nuclear + core + overlap on
the summed densities, the UHF RI-JK skeleton, for UKS the grid component with the Becke grid
shift, and the unrestricted CPHF with the RI-JK and grid responses) against central finite
differences of the unrestricted energy itself, on the water cation (doublet, 5 alpha / 4 beta);
the reference's UKS tests use the same functional lists (tests/hessian_uks/test_svwn.rs:14-15,
test_b3lyp.rs), and TPSS0 adds the spin-polarised meta-GGA.  A closed shell run through the unrestricted path must give the restricted
Hessian.

Tolerance: h = 2e-3 bohr on energies converged to 1e-13 Eh; analytic vs finite difference within
1e-4 of the largest Hessian element; unrestricted vs restricted on a closed shell within 1e-8
(the same equations, solved in different orders).  The CPHF term fills each same-atom block from
one product without symmetrising it (the reference's assembly, rscf.rs:382-388, uscf.rs:425-431);
on the cation that block is symmetric to 2e-7 of max |H| (closed shell: 1e-8), independent of
the SCF and CPHF tolerances, and Richardson-extrapolated finite differences (h = 2e-3, 1e-3) meet
both triangles within 2.3e-7, so the symmetry bound here is 1e-6."""
import numpy as np
import pytest

from tests.test_gpu_hessian_rhf import AUXB, BASIS, GEOM, NOCC, SYMS

pytestmark = pytest.mark.gpu

FUNCTIONALS = {"hf": None, "svwn": [(1.0, "LDA_X"), (1.0, "LDA_C_VWN")], "b3lyp": [(1.0, "HYB_GGA_XC_B3LYP")],
               "tpss0": [(1.0, "HYB_MGGA_XC_TPSS0")]}
Z = {"O": 8, "H": 1}
NA, NB = NOCC, NOCC - 1


def setup(R, name, spin="polarized"):
    import hesship
    from hesship.grids import Grids
    atoms = [(s, tuple(r)) for s, r in zip(SYMS, R)]
    mol, aux = hesship.CInt.from_basis(atoms, BASIS), hesship.CInt.from_basis(atoms, AUXB)
    if FUNCTIONALS[name] is None:
        return mol, aux, None, None
    xc = [(sc, hesship.LibXCFunctional.from_identifier(n, spin)) for sc, n in FUNCTIONALS[name]]
    return mol, aux, xc, Grids([Z[s] for s in SYMS], R, nrad=32, nang=110)


def scf(R, name, na=NA, nb=NB):
    import hesship
    mol, aux, xc, grids = setup(R, name)
    if xc is None:
        return hesship.UHF(mol, aux, na, nb, conv_tol=1e-13, max_cycle=200).run()
    return hesship.UKS(mol, aux, na, nb, xc, grids, conv_tol=1e-13, max_cycle=200).run()


def hessian(s, cfg=None):
    import hesship
    drv = hesship.UHessSCF(s.mol, s.aux, s.mo_coeff, s.mo_occ, s.mo_energy, xc=s.xc, grids=s.grids,
                           config=cfg or dict(cphf_tol=1e-12, cphf_max_cycle=100))
    return drv, drv.make_hess().cpu().numpy()


def square(de):
    natm = de.shape[2]
    return de.transpose(2, 0, 3, 1).reshape(3 * natm, 3 * natm)


@pytest.mark.parametrize("name", ["hf", "svwn", "b3lyp", "tpss0"])
def test_uhf_hessian_against_finite_differences(dev, name):
    s = scf(GEOM, name)
    for e, n in zip(s.mo_energy, s.nocc):
        e = e.cpu().numpy()
        assert e[n - 1] < e[n]
    drv, de = hessian(s)
    H = square(de)
    assert np.abs(H - H.T).max() < 1e-6 * np.abs(H).max()
    assert np.abs(de.sum(axis=3)).max() < 1e-6 * np.abs(de).max()
    h = 2e-3
    n = H.shape[0]
    e0 = s.e_tot
    cache = {}

    def E(*steps):
        key = tuple(sorted(steps))
        if key not in cache:
            R = GEOM.copy()
            for k, sgn in steps:
                R[k // 3, k % 3] += sgn * h
            cache[key] = scf(R, name).e_tot
        return cache[key]
    fd = np.zeros((n, n))
    for k in range(n):
        fd[k, k] = (E((k, 1)) - 2 * e0 + E((k, -1))) / h ** 2
        for l in range(k):
            fd[k, l] = fd[l, k] = (E((k, 1), (l, 1)) - E((k, 1), (l, -1)) - E((k, -1), (l, 1))
                                   + E((k, -1), (l, -1))) / (4 * h * h)
    err = np.abs(H - fd).max() / np.abs(H).max()
    parts = {k: float(np.abs(v.cpu().numpy()).max()) for k, v in drv.result.items()}
    print(f"[uhf hessian {name}] analytic vs finite difference: {err:.2e} of max |H| = {np.abs(H).max():.4f}; "
          f"SCF cycles {s.cycles}, E = {e0:.10f}; part maxima {parts}", flush=True)
    assert err < 1e-4, (err, parts)


@pytest.mark.parametrize("name", ["hf", "svwn"])
def test_closed_shell_unrestricted_equals_restricted(dev, name):
    import hesship
    u = scf(GEOM, name, NOCC, NOCC)
    _, de_u = hessian(u)
    mol, aux, xc, grids = setup(GEOM, name, "unpolarized")
    if xc is None:
        r = hesship.RHF(mol, aux, NOCC, conv_tol=1e-13, max_cycle=200).run()
    else:
        r = hesship.RKS(mol, aux, NOCC, xc, grids, conv_tol=1e-13, max_cycle=200).run()
    assert abs(u.e_tot - r.e_tot) < 1e-10
    drv = hesship.RHessSCF(mol, aux, r.mo_coeff, r.mo_occ, r.mo_energy, xc=r.xc, grids=r.grids,
                           config=dict(cphf_tol=1e-12, cphf_max_cycle=100))
    de_r = drv.make_hess().cpu().numpy()
    err = np.abs(de_u - de_r).max() / np.abs(de_r).max()
    print(f"[uhf closed shell {name}] |H_U - H_R| / max |H_R| = {err:.2e}", flush=True)
    assert err < 1e-8
