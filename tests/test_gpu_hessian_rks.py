"""End to end on the device for Kohn-Sham: density-fitted RKS (hh_scf_rks) on a partitioned
Becke grid and the whole analytic RKS Hessian (nuclear + core + RI-JK skeleton with the hybrid
K fraction + the grid component with the Becke grid shift + overlap + CPHF with the CPKS grid
response; RHessSCF::make_hess, rscf.rs:481-505, with hess_rks_becke.rs) against central finite
differences of the RKS energy itself, the grid moving with the atoms (as the Becke-shift terms
assume), for SVWN (LDA), B3LYP (hybrid GGA) and TPSS0 (hybrid meta-GGA), the functional lists of
the reference's tests (tests/hessian_uks/test_svwn.rs:14-15, tests/hessian_rks/test_b3lyp.rs:14,
test_tpss0.rs:28).

Tolerance: h = 2e-3 bohr on energies converged to 1e-13 Eh; analytic vs finite difference within
1e-4 of the largest Hessian element (the grid quadrature is the same at every displaced geometry,
so it adds no noise; the meta-GGA's tau clamp is met only in the density tails)."""
import numpy as np
import pytest

from tests.test_gpu_hessian_rhf import AUXB, BASIS, GEOM, NOCC, SYMS

pytestmark = pytest.mark.gpu

FUNCTIONALS = {"svwn": [(1.0, "LDA_X"), (1.0, "LDA_C_VWN")], "b3lyp": [(1.0, "HYB_GGA_XC_B3LYP")],
               "tpss0": [(1.0, "HYB_MGGA_XC_TPSS0")]}
Z = {"O": 8, "H": 1}


def setup(R, name):
    import hesship
    from hesship.grids import Grids
    atoms = [(s, tuple(r)) for s, r in zip(SYMS, R)]
    mol, aux = hesship.CInt.from_basis(atoms, BASIS), hesship.CInt.from_basis(atoms, AUXB)
    xc = [(sc, hesship.LibXCFunctional.from_identifier(n)) for sc, n in FUNCTIONALS[name]]
    grids = Grids([Z[s] for s in SYMS], R, nrad=32, nang=110)
    return mol, aux, xc, grids


def energy(R, name):
    import hesship
    mol, aux, xc, grids = setup(R, name)
    return hesship.RKS(mol, aux, NOCC, xc, grids, conv_tol=1e-13, max_cycle=200).run().e_tot


@pytest.mark.parametrize("name", ["svwn", "b3lyp", "tpss0"])
def test_rks_hessian_against_finite_differences(dev, name):
    import hesship
    mol, aux, xc, grids = setup(GEOM, name)
    scf = hesship.RKS(mol, aux, NOCC, xc, grids, conv_tol=1e-13, max_cycle=200).run()
    e = scf.mo_energy.cpu().numpy()
    assert e[NOCC - 1] < e[NOCC]
    drv = hesship.RHessSCF(mol, aux, scf.mo_coeff, scf.mo_occ, scf.mo_energy, xc=scf.xc, grids=scf.grids,
                           config=dict(cphf_tol=1e-12, cphf_max_cycle=100))
    de = drv.make_hess().cpu().numpy()
    natm = len(SYMS)
    H = de.transpose(2, 0, 3, 1).reshape(3 * natm, 3 * natm)
    assert np.abs(H - H.T).max() < 1e-7 * np.abs(H).max()
    assert np.abs(de.sum(axis=3)).max() < 1e-6 * np.abs(de).max()
    h = 2e-3
    n = 3 * natm
    e0 = scf.e_tot
    cache = {}

    def E(*steps):
        key = tuple(sorted(steps))
        if key not in cache:
            R = GEOM.copy()
            for k, sgn in steps:
                R[k // 3, k % 3] += sgn * h
            cache[key] = energy(R, name)
        return cache[key]
    fd = np.zeros((n, n))
    for k in range(n):
        fd[k, k] = (E((k, 1)) - 2 * e0 + E((k, -1))) / h ** 2
        for l in range(k):
            fd[k, l] = fd[l, k] = (E((k, 1), (l, 1)) - E((k, 1), (l, -1)) - E((k, -1), (l, 1))
                                   + E((k, -1), (l, -1))) / (4 * h * h)
    err = np.abs(H - fd).max() / np.abs(H).max()
    parts = {k: float(np.abs(v.cpu().numpy()).max()) for k, v in drv.result.items()}
    print(f"[rks hessian {name}] analytic vs finite difference: {err:.2e} of max |H| = {np.abs(H).max():.4f}; "
          f"SCF cycles {scf.cycles}, E = {e0:.10f}; part maxima {parts}", flush=True)
    assert err < 1e-4, (err, parts)
