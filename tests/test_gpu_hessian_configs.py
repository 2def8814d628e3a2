"""The whole Hessian on the molecules of the configurations (basis sets synthetic: no basis-set
files exist here, so the reference's fingerprints cannot be compared):

* c1-shaped: NH3 RHF with a cc-pVDZ-sized basis (nao 29, nocc 5, natm 4, nprop 12), against
  central finite differences of the device SCF energy, and its six harmonic frequencies;
* c3: the seeded (H2O)_8 cluster with B3LYP on the 75 x 302 Becke grid (nao 192, naux 632,
  543 600 points, nprop 72): symmetric and translation invariant within 2e-6 of the largest
  element (the CPHF residual 1e-9 leaves ~3e-7), and the mixed-precision CPHF (BF16x3 with FP64
  refinement) agrees with the FP64 CPHF within 1e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NH3_BASIS = {"N": [(0, [[9046.0, 0.0007], [1357.0, 0.0054], [309.3, 0.0274], [87.73, 0.1032], [28.56, 0.2782],
                        [10.21, 0.4485], [3.838, 0.2786]]), (0, [[0.7466, 1.0]]), (0, [[0.2248, 1.0]]),
                   (1, [[13.55, 0.0399], [2.917, 0.2172], [0.7973, 0.5103]]), (1, [[0.2185, 1.0]]), (2, [[0.817, 1.0]])],
             "H": [(0, [[13.01, 0.0197], [1.962, 0.138], [0.4446, 0.478]]), (0, [[0.122, 1.0]]), (1, [[0.727, 1.0]])]}
NH3_AUX = {"N": [(0, [[12.0, 1.0]]), (0, [[3.0, 1.0]]), (0, [[0.8, 1.0]]), (1, [[4.0, 1.0]]), (1, [[1.0, 1.0]]),
                 (2, [[2.5, 1.0]]), (2, [[0.7, 1.0]]), (3, [[1.4, 1.0]])],
           "H": [(0, [[2.5, 1.0]]), (0, [[0.6, 1.0]]), (1, [[1.2, 1.0]]), (2, [[0.9, 1.0]])]}
NH3_GEOM = np.array([[0.0, 0.0, 0.13], [0.0, 1.77, -0.60], [1.53, -0.88, -0.60], [-1.53, -0.88, -0.60]])
NH3_SYMS = ["N", "H", "H", "H"]


def nh3(R):
    import hesship
    atoms = [(s, tuple(r)) for s, r in zip(NH3_SYMS, R)]
    return hesship.CInt.from_basis(atoms, NH3_BASIS), hesship.CInt.from_basis(atoms, NH3_AUX)


def test_c1_nh3_rhf_hessian(dev):
    import hesship
    from hesship import vib
    mol, aux = nh3(NH3_GEOM)
    assert mol.nao == 29
    scf = hesship.RHF(mol, aux, 5, conv_tol=1e-13, max_cycle=200).run()
    drv = hesship.RHessSCF(mol, aux, scf.mo_coeff, scf.mo_occ, scf.mo_energy, config=dict(cphf_tol=1e-12))
    de = drv.make_hess().cpu().numpy()
    n = 12
    H = de.transpose(2, 0, 3, 1).reshape(n, n)
    h = 2e-3
    cache = {}

    def E(*steps):
        key = tuple(sorted(steps))
        if key not in cache:
            R = NH3_GEOM.copy()
            for k, sgn in steps:
                R[k // 3, k % 3] += sgn * h
            m, a = nh3(R)
            cache[key] = hesship.RHF(m, a, 5, conv_tol=1e-13, max_cycle=200).run().e_tot
        return cache[key]
    fd = np.zeros((n, n))
    for k in range(n):
        fd[k, k] = (E((k, 1)) - 2 * scf.e_tot + E((k, -1))) / h ** 2
        for l in range(k):
            fd[k, l] = fd[l, k] = (E((k, 1), (l, 1)) - E((k, 1), (l, -1)) - E((k, -1), (l, 1))
                                   + E((k, -1), (l, -1))) / (4 * h * h)
    err = np.abs(H - fd).max() / np.abs(H).max()
    r = vib.harmonic_analysis(0.5 * (H + H.T), NH3_GEOM.T, np.array([14.00307401, 1.00782503, 1.00782503, 1.00782503]))
    freq = np.where(r.imag[r.vib_indices], -1.0, 1.0) * r.omega[r.vib_indices]
    print(f"[c1 nh3] analytic vs finite difference {err:.2e}; E = {scf.e_tot:.10f}; frequencies {np.round(freq, 1)}",
          flush=True)
    assert err < 2e-5
    assert len(freq) == 6


def test_c3_water8_b3lyp_hessian(dev):
    import hesship
    from hesship import synth
    from hesship.grids import Grids
    from bench import JK_LIKE, SVP_LIKE
    z, R = synth.water_cluster(8, box_ang=7.0, seed=0x5EED0008)
    sym = {8: "O", 1: "H"}
    atoms = [(sym[int(c)], tuple(x)) for c, x in zip(z, R)]
    mol, aux = hesship.CInt.from_basis(atoms, SVP_LIKE), hesship.CInt.from_basis(atoms, JK_LIKE)
    assert (mol.nao, aux.nao) == (192, 632)
    xc = [(1.0, hesship.LibXCFunctional.from_identifier("HYB_GGA_XC_B3LYP"))]
    grids = Grids(z, R, nrad=75, nang=302)
    scf = hesship.RKS(mol, aux, 40, xc, grids, conv_tol=1e-11, max_cycle=200).run()
    assert grids.coords.shape[0] == 543600
    out = {}
    for prec, tol in (("fp64", 1e-9), ("bf16x3", 1e-9)):
        drv = hesship.RHessSCF(mol, aux, scf.mo_coeff, scf.mo_occ, scf.mo_energy, xc=scf.xc, grids=scf.grids,
                               config=dict(cphf_tol=tol))
        out[prec] = drv.make_hess(precision=prec).cpu().numpy()
    de = out["fp64"]
    natm = len(atoms)
    H = de.transpose(2, 0, 3, 1).reshape(3 * natm, 3 * natm)
    sc = np.abs(H).max()
    asym = np.abs(H - H.T).max() / sc
    trans = np.abs(de.sum(axis=3)).max() / sc
    diff = np.abs(out["bf16x3"] - de).max() / sc
    print(f"[c3 (H2O)8 b3lyp] max |H| {sc:.4f}; asymmetry {asym:.1e}, translation {trans:.1e}, "
          f"mixed vs FP64 {diff:.2e}", flush=True)
    # at the CPHF tolerance 1e-9 the asymmetry is ~3e-7 of the largest element (north-star accuracy 1e-6)
    assert asym < 2e-6 and trans < 2e-6
    assert diff < 1e-6
