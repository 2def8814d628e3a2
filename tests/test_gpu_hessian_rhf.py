"""End to end on the device: density-fitted RHF (hh_scf_rhf) and the whole analytic RHF Hessian
(nuclear repulsion + core Hamiltonian + RI-JK skeleton + overlap + CPHF, RHessSCF::make_hess,
rscf.rs:481-505) of a water molecule in a synthetic basis, against central finite differences
of the SCF energy itself: the analytic Hessian is the second derivative of exactly that energy
(the auxiliary functions move with the atoms, as in the reference's RI-JK component), so a wrong
factor, sign or index in any term shows.  Also: translation invariance, symmetry, and the
harmonic frequencies of the result.

Tolerance: the finite differences use h = 2e-3 bohr on energies converged to 1e-13 Eh (noise
~1e-8, truncation ~h^2 / 12 E'''' ~ 1e-6); analytic vs finite difference within 2e-5 of the
largest Hessian element."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASIS = {"O": [(0, [[130.7, 0.154], [23.81, 0.535], [6.44, 0.445]]), (0, [[0.82, 1.0]]),
               (1, [[5.03, 0.156], [1.17, 0.607]]), (1, [[0.38, 1.0]])],
         "H": [(0, [[3.42, 0.155], [0.62, 0.535]]), (1, [[0.9, 1.0]])]}
AUXB = {"O": [(0, [[9.0, 1.0]]), (0, [[2.2, 1.0]]), (0, [[0.6, 1.0]]), (1, [[3.0, 1.0]]), (1, [[0.8, 1.0]]),
              (2, [[1.5, 1.0]])],
        "H": [(0, [[2.5, 1.0]]), (0, [[0.6, 1.0]]), (1, [[1.0, 1.0]])]}
GEOM = np.array([[0.0, 0.02, 0.11], [0.0, 1.45, -0.93], [0.03, -1.47, -0.88]])
SYMS = ["O", "H", "H"]
NOCC = 5


def molecule(R):
    import hesship
    atoms = [(s, tuple(r)) for s, r in zip(SYMS, R)]
    return hesship.CInt.from_basis(atoms, BASIS), hesship.CInt.from_basis(atoms, AUXB)


def energy(R):
    import hesship
    mol, aux = molecule(R)
    return hesship.RHF(mol, aux, NOCC, conv_tol=1e-13, max_cycle=200).run().e_tot


def test_rhf_hessian_against_finite_differences(dev):
    import hesship
    mol, aux = molecule(GEOM)
    scf = hesship.RHF(mol, aux, NOCC, conv_tol=1e-13, max_cycle=200).run()
    e = scf.mo_energy.cpu().numpy()
    assert scf.cycles < 60 and np.all(np.diff(e) >= 0) and e[NOCC - 1] < e[NOCC]
    drv = hesship.RHessSCF(mol, aux, scf.mo_coeff, scf.mo_occ, scf.mo_energy,
                           config=dict(cphf_tol=1e-12, cphf_max_cycle=100))
    de = drv.make_hess().cpu().numpy()
    natm = len(SYMS)
    H = de.transpose(2, 0, 3, 1).reshape(3 * natm, 3 * natm)   # [(A, i), (B, j)] = de[i, j, A, B]
    assert np.abs(H - H.T).max() < 1e-8 * np.abs(H).max()
    assert np.abs(de.sum(axis=3)).max() < 1e-7 * np.abs(de).max()   # translation invariance
    # finite differences of the SCF energy
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
            cache[key] = energy(R)
        return cache[key]
    fd = np.zeros((n, n))
    for k in range(n):
        fd[k, k] = (E((k, 1)) - 2 * e0 + E((k, -1))) / h ** 2
        for l in range(k):
            fd[k, l] = fd[l, k] = (E((k, 1), (l, 1)) - E((k, 1), (l, -1)) - E((k, -1), (l, 1))
                                   + E((k, -1), (l, -1))) / (4 * h * h)
    err = np.abs(H - fd).max() / np.abs(H).max()
    parts = {k: float(np.abs(v.cpu().numpy()).max()) for k, v in drv.result.items()}
    print(f"[rhf hessian] analytic vs finite difference: {err:.2e} of max |H| = {np.abs(H).max():.4f}; "
          f"SCF cycles {scf.cycles}, E = {e0:.10f}; part maxima {parts}", flush=True)
    assert err < 2e-5, (err, parts)
    # the harmonic analysis of the analytic Hessian (vib.rs:340): three real vibrations
    from hesship import vib
    r = vib.harmonic_analysis(H, GEOM.T, np.array([15.99491462, 1.00782503, 1.00782503]))
    idx = r.vib_indices
    assert len(idx) == 3 and not np.any(r.imag[idx]) and np.all(r.omega[idx] > 500), r.omega


def test_scf_energy_is_stationary_and_reproducible(dev):
    import hesship
    mol, aux = molecule(GEOM)
    a = hesship.RHF(mol, aux, NOCC, conv_tol=1e-13).run()
    b = hesship.RHF(mol, aux, NOCC, conv_tol=1e-13).run()
    assert abs(a.e_tot - b.e_tot) < 1e-12
    # a translated molecule has the same energy
    mol2, aux2 = molecule(GEOM + np.array([0.3, -0.2, 0.5]))
    c = hesship.RHF(mol2, aux2, NOCC, conv_tol=1e-13).run()
    assert abs(a.e_tot - c.e_tot) < 1e-10
    # orbital energies: the occupied block below the virtual block; C^T S C = I
    S = mol.intor("int1e_ovlp").cpu().numpy()
    Cm = a.mo_coeff.cpu().numpy()
    assert np.abs(Cm.T @ S @ Cm - np.eye(mol.nao)).max() < 1e-10
