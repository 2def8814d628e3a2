"""Benchmark: RI-JK response (J + K, all perturbations) on the synthetic workload, one MI355X.

Workload "RI-JK response kernel on synthetic tensors": naux=2400, nao=400, nocc=40, nvir=360,
nprop=90 (30-atom equivalent), seeded as in hesship/synth.py.  One step = one full
get_response_bra call (J and K for all 90 perturbed densities), inputs resident in HBM.

Prints ONE JSON line.  `value` = RI-K TFLOP/s of the production precision, computed from the
algorithmic count F_K = 2 nao^2 naux nocc + 2 nao naux nocc^2 + nprop (2 nao^2 naux nocc +
6 nao naux nocc^2).  The plain FP64 path is timed beside it, and the production result's error
against the FP64 result is reported.

  python bench.py [--steps K] [--warmup W] [--precision bf16x3|fp32|fp64] [--gpus 1]

`--workload` selects a measurement of one of the other hot-path pieces instead (one JSON line
each, same fields): `numint` (RKS grid response at the (H2O)_8 B3LYP sizes on synthetic AO
values), `dft_setup` (DFT skeleton setup passes at the same sizes, FP64), `dft_becke` (the same
setup with the Becke grid shift on an atom-grouped grid, FP64), `skeleton` (RI-JK skeleton
Hessian and first-derivative terms on synthetic integrals, FP64), `cphf_batch` (64 uneven
alkane-sized RI-JK CPHF solves, half of them doublet cations (UHF), on one stream: plain FP64 vs
mixed precision with HIP-graph replay of the response), `xc` (B3LYP exc, vxc_eff and fxc_eff on a
synthetic 543 600-point density, closed shell and spin polarised, FP64).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "rstsr-showcase-hessian_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "CPHF+Hessian wall s & RI-K TFLOP/s: mixed+learned vs plain FP64, 1 MI355X"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}  # MI355X dense matrix peaks
# HBM bytes per launch of the hot kernel, measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per
# dispatch) and scaled by the same counters on a 1 GiB copy16_kernel copy (factors 1.99994 and 1.00000);
# profiles/nt_bf16x3_kernel_ri_k_nao400_naux2400_nocc40_nprop90_hbm_traffic.txt.  Keyed by
# (precision, nao, naux, nocc, nprop, BF16X3 tile configuration); other workloads report null.
MEASURED_TRAFFIC = {("bf16x3", 400, 2400, 40, 90, 0): 13881227.0 * 1024 * 1.99994 + 91504.8 * 1024 * 1.00000,
                    ("bf16x3", 400, 2400, 40, 90, 2): 7991991.2 * 1024 * 1.99994 + 91596.1 * 1024 * 1.00000,
                    # fused term-1 kernel: profiles/kfused_kernel_ri_k_nao400_naux2400_nocc40_nprop90_counters.txt
                    ("bf16x3", 400, 2400, 40, 90, "fused"): 1720391.2 * 1024 * 1.99994 + 152100.0 * 1024 * 1.00000}
# rocprofv3 --pmc counters of the same launches (same profile): SQ_VALU_MFMA_BUSY_CYCLES over
# GRBM_GUI_ACTIVE x 1024 SIMDs, SQ_LDS_IDX_ACTIVE over GRBM_GUI_ACTIVE x 256 CUs, SQ_LDS_BANK_CONFLICT
MEASURED_COUNTERS = {("bf16x3", 400, 2400, 40, 90, "fused"): {"mfma_busy": 0.559, "lds_busy": 0.186,
                                                            "lds_bank_conflict_cycles": 0,
                                                            "engine_clock_ghz": 2.03}}
HOT_KERNEL = {"fp64": "nt_f64_kernel<1,8,13,1>", "fp32": "gemm_kernel<1,float,double>",
              "bf16x3": "nt_bf16x3_kernel<1,8,13,2,true,13,2>"}
# BF16X3 tile configurations of the NT kernel (the autotuner's candidates), by index
BF16X3_CONFIGS = {0: "nt_bf16x3_kernel<1,8,13,2,true,13,2> (208x256 tile)",
                  1: "nt_bf16x3_kernel<2,4,7,4,true,13,1> (208x256 tile)",
                  2: "nt_bf16x3_kernel<2,4,13,2,true,25,2,swizzled> (400x128 tile)"}
SPLIT_RULES = {0: "about one workgroup per CU", -1: "last wave filled"}


def kfused_supported(nao, nocc):
    """The fused BF16x3 K path's range (rijk_kfused.hip: kfused_supported)."""
    return 8 <= nocc <= 80 and -(-nao // 32) <= 13 and 2 + 30 // nocc <= (5 if nocc <= 16 else 3 if nocc <= 32 else 2)


def kfused_name(nao, nocc):
    nb = -(-nao // 32)
    nmb = 4 if nb <= 4 else 8 if nb <= 8 else 10 if nb <= 10 else 12 if nb <= 12 else 13
    nh = -(-nocc // 16)
    na = {1: 5, 2: 3}.get(nh, 2)
    return f"kfused_kernel<{nmb},{nh},{na}> (all {32 * nmb} rows x 128 columns per workgroup, 32x32x16 MFMA)"


def measured_mfma_peak(precision, kind=None):
    """Issue-rate ceiling of the MFMA instruction the hot kernel uses, measured on this device
    (hh_util_mfma_peak: independent accumulator chains, no memory traffic), TFLOP/s."""
    import ctypes as C
    import torch
    from hesship import _lib
    if kind is None:
        kind = 1 if precision == "fp64" else 0
    iters = {0: 500, 1: 20000, 2: 250}[kind]
    out = C.c_double()
    _lib.check(_lib.lib().hh_util_mfma_peak(kind, iters, C.byref(out), torch.cuda.current_stream().cuda_stream))
    return out.value


def hot_choice(nao, naux, nocc, nprop):
    """The autotuner's choice for the BF16X3 hot contraction (M nao, N nocc nprop, K kpad naux)."""
    import ctypes as C
    from hesship import _lib
    cfg, ks = C.c_int32(), C.c_int32()
    kpad = -(-nao // 32) * 32
    _lib.check(_lib.lib().hh_util_gemm_nt_bf16x3_choice(nao, nocc * nprop, kpad * naux, C.byref(cfg), C.byref(ks)))
    return cfg.value, ks.value
SIZES = {
    "full": dict(nao=400, naux=2400, nocc=40, nprop=90),
    "small": dict(nao=96, naux=480, nocc=12, nprop=18),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_k(nao, naux, nocc, nprop):
    return 2 * nao**2 * naux * nocc + 2 * nao * naux * nocc**2 + nprop * (2 * nao**2 * naux * nocc + 6 * nao * naux * nocc**2)


def flops_j(nao, naux, nocc, nprop):
    ntp = nao * (nao + 1) // 2
    return 4 * ntp * naux * nprop + 4 * nao**2 * nocc * nprop


def cpu_baseline(cderi, C, occ, bra, budget_s=12.0, nbatch=72):
    """The oracle's restatement of the reference K loop (hess_r.rs:144-183: aux batches of 72,
    serial per-perturbation GEMMs), timed on host cores over a bounded number of aux batches."""
    import numpy as np
    from oracle import rijk as orc
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nao = C.shape[0]
    naux = cderi.shape[1]
    mo = C[:, np.asarray(occ) > 0]
    nocc = mo.shape[1]
    nprop = bra.shape[2]
    orc.k_batch(cderi[:, :8], mo, bra[:, :, :2])  # warm-up
    done, t0 = 0, time.perf_counter()
    while done < naux:
        p1 = min(naux, done + nbatch)
        orc.k_batch(cderi[:, done:p1], mo, bra)
        done = p1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    # per-aux share of F_K (every term of F_K is linear in naux)
    fk = flops_k(nao, done, nocc, nprop)
    return dict(value=round(fk / dt / 1e12, 5), unit="TFLOP/s", cores=int(threads), kind="port",
                sample=f"RI-K loop of the reference restated in NumPy FP64: {done} of {naux} aux "
                       f"({-(-done // nbatch)} batches of {nbatch}), all {nprop} perturbations, {dt:.1f} s")


def cphf_case(obj, d_C, occ, nao, nocc, nprop, dev, config=None, rhs_seed=0x5EED0004):
    """Synthetic CPHF system on the benchmark's RI-JK tensors: orbital energies and first-order
    Fock/overlap derivatives drawn from the same seeded generator (seed 0x5EED0004; rhs_seed
    replaces the first-order derivatives only); the RI-JK scale is set so the dimensionless
    response has norm ~0.3 of its input, as in a typical closed-shell CPHF."""
    import numpy as np
    import torch
    import hesship
    from hesship import synth
    z = synth.normal(0x5EED0004, 0, nao + 2 * nao * nocc * nprop)
    e = np.sort(np.concatenate([-1.4 + 1.1 * np.abs(np.tanh(z[:nocc])), 0.2 + 2.5 * np.abs(np.tanh(z[nocc:nao]))]))
    if rhs_seed != 0x5EED0004:
        z = np.concatenate([z[:nao], synth.normal(rhs_seed, 0, 2 * nao * nocc * nprop)])
    f1 = z[nao:nao + nao * nocc * nprop].reshape(nprop, nocc, nao).T.copy()
    s1 = 0.1 * z[nao + nao * nocc * nprop:].reshape(nprop, nocc, nao).T.copy()
    drv = hesship.RHessCPHF(d_C, occ, e, dict({"cphf_tol": 1e-8}, **(config or {})))
    drv.add_el(obj)
    v = torch.from_numpy(np.random.default_rng(1).standard_normal((nao, nocc, 4))).to(dev)
    ratio = (drv.response_dimless_cphf(v).norm() / v.norm()).item()
    return drv, e, torch.from_numpy(f1).to(dev), torch.from_numpy(s1).to(dev), ratio


def response_spectrum(drv, nao, nocc, dev, which=("SR", "LR")):
    """Lowest and highest eigenvalue of the dimensionless response operator A on the vir-occ
    block(s) (I + A is the CPHF matrix; the occupied rows of the response are zero): ARPACK (scipy
    eigs) with one FP64 response per matvec.  A = D^-1 K with D the orbital-energy gaps and K
    symmetric, so the spectrum is real.  `nocc`: an int (RHF driver) or one count per spin (UHF)."""
    import numpy as np
    import torch
    from scipy.sparse.linalg import LinearOperator, eigs
    noc = [nocc] if isinstance(nocc, (int, np.integer)) else list(nocc)
    sizes = [(nao - no) * no for no in noc]
    n = sum(sizes)

    def mv(v):
        v = np.real(v)
        xs, p = [], 0
        for no, sz in zip(noc, sizes):
            x = np.zeros((nao, no, 1))
            x[no:, :, 0] = v[p:p + sz].reshape(nao - no, no)
            xs.append(torch.from_numpy(x).to(dev))
            p += sz
        ys = drv.response_dimless_cphf(xs[0] if len(noc) == 1 else xs)
        ys = [ys] if len(noc) == 1 else ys
        return np.concatenate([y[no:, :, 0].cpu().numpy().ravel() for y, no in zip(ys, noc)])

    op = LinearOperator((n, n), matvec=mv, dtype=np.float64)
    v0 = np.random.default_rng(7).standard_normal(n)
    out = []
    for w in which:
        ev = eigs(op, k=2, which=w, return_eigenvectors=False, tol=1e-6, ncv=min(24, n - 1), v0=v0)
        out.append(float(np.real(ev).min() if w == "SR" else np.real(ev).max()))
    return tuple(out)


def resident_leg(d_cderi, d_C, occ, d_bra, r64, nao, nocc, nprop, steps):
    """The resident RHF response (rijk_resident.hip): M [nao nocc, nao nocc] built from the same
    tensors (timed: the first call after make_response_preparation), then `steps` calls, each one
    product that streams M once.  Roofline: HBM, M's bytes per call / call time."""
    import torch
    import hesship
    out = {}
    R = nao * nocc
    kpad = -(-R // 32) * 32
    scale = r64.abs().max().item()
    for prec in ("bf16x3", "fp64"):
        obj = hesship.RHessRIJK(d_cderi, 1.0, 1.0)
        obj.set_response_path("resident")
        obj.make_response_preparation(d_C, occ)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = obj.get_response_bra(d_bra, precision=prec)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            r = obj.get_response_bra(d_bra, precision=prec)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        mbytes = R * kpad * (4 if prec == "bf16x3" else 8)
        gbs = mbytes / (ms * 1e-3) / 1e9
        out[prec] = {"build_plus_first_call_s": round(build_s, 4), "ms_per_call": round(ms, 4),
                     "matrix_bytes": mbytes, "achieved_GBps": round(gbs, 1), "hbm_frac": round(gbs / 8000.0, 4),
                     "build_flops": 4.0 * R * R * (d_cderi.shape[1] if d_cderi.dim() > 1 else 0),
                     "rel_err_vs_fp64_direct": (r - r64).abs().max().item() / scale}
        del obj
        torch.cuda.empty_cache()
    out["note"] = ("resp = M bra with M = 4 sj sum_P X_P (x) X_P - sk sum_P (B_P (x) O_P + X_P (x) X_P) built once per "
                   "orbital set (two products of 2 (nao nocc)^2 naux flop); per call 2 (nao nocc)^2 nprop flop")
    return out


def model_frequency_shift(dde):
    """Largest change (cm^-1) of the harmonic frequencies of a model (H2O)_n Hessian when the
    CPHF Hessian difference dde [3, 3, natm, natm] (mixed - FP64, Eh/bohr^2) is added to it.  The
    model: n waters on a 3 x 3 grid of layers 5.5 bohr apart (O-H 1.81 bohr, 104.5 degrees),
    central springs along every atom pair (0.5 Eh/bohr^2 for the O-H bonds, 0.05 exp(-r / 5 bohr)
    otherwise: a rigid network, translation and rotation invariant at this geometry, lowest
    vibration 171 cm^-1 for 30 atoms); analysed with the library's harmonic analysis
    (hh_vib_harmonic_analysis, TR projected), real vibrational modes only."""
    import numpy as np
    from hesship import vib
    natm = dde.shape[-1]
    nw = natm // 3
    geom = []
    for w in range(nw):
        o = 5.5 * np.array([w % 3, (w // 3) % 3, w // 9], dtype=float)
        a = np.deg2rad(104.5) / 2
        geom += [o, o + 1.81 * np.array([np.sin(a), np.cos(a), 0.0]), o + 1.81 * np.array([-np.sin(a), np.cos(a), 0.0])]
    geom = np.array(geom[:natm]).T.copy()
    mass = np.array(([15.99491, 1.00783, 1.00783] * nw)[:natm])
    h = np.zeros((3 * natm, 3 * natm))
    for a in range(natm):
        for b in range(a + 1, natm):
            d = geom[:, b] - geom[:, a]
            r = np.linalg.norm(d)
            k = (0.5 if r < 2.2 else 0.05 * np.exp(-r / 5.0)) * np.outer(d, d) / r**2
            for i, j, sg in ((a, a, 1), (b, b, 1), (a, b, -1), (b, a, -1)):
                h[3 * i:3 * i + 3, 3 * j:3 * j + 3] += sg * k
    d = dde.transpose(2, 0, 3, 1).reshape(3 * natm, 3 * natm)  # [A t, B s]
    d = 0.5 * (d + d.T)
    f0 = vib.harmonic_analysis(h, geom, mass)
    f1 = vib.harmonic_analysis(h + d, geom, mass)
    v0 = np.sort([f0.omega[i] for i in f0.vib_indices if not f0.imag[i]])
    v1 = np.sort([f1.omega[i] for i in f1.vib_indices if not f1.imag[i]])
    n = min(len(v0), len(v1))
    return float(np.abs(v1[:n] - v0[:n]).max())


def run_cphf(drv, f1, s1, prec, prepare=None):
    """rhs -> Krylov -> finalize, timed; `prepare` (re-preparing the RI-JK component, which for the
    resident operator includes building its response matrix on the first response) inside the
    timed region."""
    import torch
    st = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if prepare is not None:
        prepare()
    rhs = drv.compute_dimless_rhs(f1, s1)
    mo1 = drv.solve_dimless_cphf(rhs, precision=prec, stats=st)
    mo1f, moe1 = drv.finalize_cphf(f1, s1, mo1)
    torch.cuda.synchronize()
    return mo1f, moe1, time.perf_counter() - t0, st


LOWEST_A = 0.75  # the synthetic CPHF matrix I + A has smallest eigenvalue 0.25


def learned_leg(resident, d_C, occ, nao, nocc, nprop, dev, f1, s1, ta, mo1t, de_t):
    """Mixed-precision solve with the resident operator and the learned diagonal preconditioner
    (hesship.precond): the network is trained on (residual, solution) pairs of the method's own
    FP64 solve of the same operator with another right-hand side (seed 0x5EED0006), then applied
    to the benchmark's system; the timed region holds prepare_response, the features and the
    network's forward pass, the solve and finalize."""
    import torch
    from hesship import precond
    t0 = time.perf_counter()
    drv_t, _, f1t, s1t, _ = cphf_case(resident, d_C, occ, nao, nocc, nprop, dev,
                                      {"cphf_tol": 1e-10, "cphf_lindep": 1e-24, "cphf_max_cycle": 100},
                                      rhs_seed=0x5EED0006)
    rhs_t = drv_t.compute_dimless_rhs(f1t, s1t)
    sol_t = drv_t.solve_dimless_cphf(rhs_t, precision="fp64")
    pair = precond.pair_norms(resident, d_C)
    net = precond.DiagMLP(seed=11)
    xt = precond.features(drv_t, pair, rhs_t)
    tt, wt = precond.target(drv_t, rhs_t, sol_t)
    loss = net.train(xt, tt, wt, epochs=500, lr=3e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del drv_t
    drv, _, _, _, _ = cphf_case(resident, d_C, occ, nao, nocc, nprop, dev, {"cphf_tol": 1e-8})

    def solve():
        st = {}
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        resident.make_response_preparation(d_C, occ)
        rhs = drv.compute_dimless_rhs(f1, s1)
        d = precond.diag(drv, net.forward(precond.features(drv, pair, rhs)))
        x = drv.solve_dimless_cphf(rhs, precision="bf16x3", diag=d, stats=st)
        mo1f, moe1 = drv.finalize_cphf(f1, s1, x)
        torch.cuda.synchronize()
        return mo1f, moe1, time.perf_counter() - w0, st, rhs, x

    solve()  # untimed
    mo1, moe1, wall, st, rhs, x = solve()
    de = drv.get_cphf_hess(ta(f1), ta(s1), ta(mo1), ta(moe1))
    # held-out fit: the network against the least-squares diagonal of this system's solution
    te, we = precond.target(drv, rhs, x)
    xe = precond.features(drv, pair, rhs)
    held, _ = net.loss_and_grad(xe, te, we)
    fresh, _ = precond.DiagMLP(seed=11).loss_and_grad(xe, te, we)
    return {"precision": "bf16x3", "operator": "resident", "preconditioner": "learned diagonal (MLP 3-64-64-64-1)",
            "cphf_tol": 1e-8, "wall_s": round(wall, 4), "cycles": st.get("cycles"), "aop_calls": st.get("aop_calls"),
            "refinements": st.get("refinements"), "residual": st.get("residual"),
            "fp64_fallback": st.get("fp64_fallback"),
            "max_abs_mo1_err_vs_tight": (mo1 - mo1t).abs().max().item(),
            "max_abs_de_cphf_err_vs_tight": (de - de_t).abs().max().item(),
            "freq_shift_vs_tight_cm1": model_frequency_shift((de - de_t).cpu().numpy()),
            "training": {"pairs": int(xt.shape[0]), "epochs": 500, "final_loss": loss,
                         "held_out_loss": {"trained": held, "untrained": fresh}, "seconds_untimed": round(train_s, 3)}}


def cphf_legs(obj, d_cderi, d_C, occ, nao, nocc, nprop, dev):
    """The wall-time half of the metric on a synthetic CPHF system built on the benchmark tensors
    (RI-JK scaled so the dimensionless response is ~0.3 of its input, or less where that would
    leave I + A indefinite; the raw N(0,1) cderi would make I + A far from a physical CPHF).  Every leg's wall time covers prepare_response of the
    RI-JK component (X, O; for the resident operator also its response matrix), the right-hand
    side, the Krylov solve and finalize.  Legs: the plain FP64 path (the reference's per-call
    algorithm and stopping rule), FP64 asked for the mixed solves' accuracy, FP64 with the
    resident operator, and mixed precision (BF16x3 inner solves, FP64 residuals) with the per-call
    and the resident operator at residual tolerances 1e-8 and 1e-9.  Accuracy against the same
    system solved in FP64 to 1e-11."""
    import hesship
    probe, _, _, _, ratio = cphf_case(obj, d_C, occ, nao, nocc, nprop, dev)
    del probe
    # A is quadratic in cderi: scale2 multiplies its spectrum.  First |A v| ~ 0.3 |v| for random v,
    # then, if A has eigenvalues below -LOWEST_A, shrink it so that I + A stays positive definite
    # with smallest eigenvalue 1 - LOWEST_A, as for a stable closed-shell ground state.
    scale2 = 0.3 / ratio
    pre = hesship.RHessRIJK(d_cderi * scale2 ** 0.5, 1.0, 1.0)
    pre.set_response_path("resident")
    pre.make_response_preparation(d_C, occ)
    drv0 = cphf_case(pre, d_C, occ, nao, nocc, nprop, dev)[0]
    lo0, hi0 = response_spectrum(drv0, nao, nocc, dev)
    del drv0, pre
    if lo0 < -LOWEST_A:
        scale2 *= LOWEST_A / -lo0
    cd = d_cderi * scale2 ** 0.5
    direct = hesship.RHessRIJK(cd, 1.0, 1.0)
    direct.make_response_preparation(d_C, occ)
    resident = hesship.RHessRIJK(cd, 1.0, 1.0)
    resident.set_response_path("resident")
    resident.make_response_preparation(d_C, occ)
    res = {"natm_equiv": nprop // 3}
    drv1 = cphf_case(resident, d_C, occ, nao, nocc, nprop, dev)[0]
    lo1, hi1 = response_spectrum(drv1, nao, nocc, dev)
    del drv1
    res["spectrum_A"] = {"min": round(lo1, 4), "max": round(hi1, 4),
                         "at_ratio_0.3": [round(lo0, 4), round(hi0, 4)],
                         "note": "eigenvalues of the dimensionless response A (I + A is solved); the cderi "
                                 "scale puts min(A) at -%.2f when |A v| ~ 0.3 |v| would make I + A "
                                 "indefinite" % LOWEST_A}
    ta = lambda x: x.reshape(*x.shape[:-1], nprop // 3, 3).transpose(-2, -1)  # p = t + 3 A
    tight, e, f1, s1, ratio2 = cphf_case(direct, d_C, occ, nao, nocc, nprop, dev,
                                         {"cphf_tol": 1e-11, "cphf_lindep": 1e-24, "cphf_max_cycle": 100})
    res["response_ratio"] = round(ratio2, 4)
    mo1t, moe1t, _, stt = run_cphf(tight, f1, s1, "fp64")
    de_t = tight.get_cphf_hess(ta(f1), ta(s1), ta(mo1t), ta(moe1t))
    res["accuracy_reference"] = {"cphf_tol": 1e-11, "cphf_lindep": 1e-24, "cycles": stt.get("cycles")}
    del tight
    legs = [  # name, component, precision, config
        ("fp64", direct, "fp64", {"cphf_tol": 1e-8}),
        ("fp64_matched", direct, "fp64", {"cphf_tol": 1e-9, "cphf_lindep": 1e-24, "cphf_max_cycle": 100}),
        ("fp64_resident", resident, "fp64", {"cphf_tol": 1e-8}),
        ("mixed_direct", direct, "bf16x3", {"cphf_tol": 1e-8}),
        ("mixed_resident", resident, "bf16x3", {"cphf_tol": 1e-8}),
        ("mixed_resident_tol1e-9", resident, "bf16x3", {"cphf_tol": 1e-9}),
    ]
    for name, comp, prec, cfg in legs:
        drv, _, _, _, _ = cphf_case(comp, d_C, occ, nao, nocc, nprop, dev, cfg)
        drv.set_verbose(name.startswith("mixed"))
        prep = lambda c=comp: c.make_response_preparation(d_C, occ)
        run_cphf(drv, f1, s1, prec, prep)  # untimed: products of new shapes are tuned on first use
        mo1, moe1, wall, st = run_cphf(drv, f1, s1, prec, prep)
        de = drv.get_cphf_hess(ta(f1), ta(s1), ta(mo1), ta(moe1))
        res[name] = {"precision": prec, "operator": "resident" if comp is resident else "per call",
                     "cphf_tol": cfg["cphf_tol"], "wall_s": round(wall, 4), "cycles": st.get("cycles"),
                     "aop_calls": st.get("aop_calls"), "refinements": st.get("refinements"),
                     "residual": st.get("residual"), "fp64_fallback": st.get("fp64_fallback"),
                     "max_abs_mo1_err_vs_tight": (mo1 - mo1t).abs().max().item(),
                     "max_abs_de_cphf_err_vs_tight": (de - de_t).abs().max().item(),
                     "freq_shift_vs_tight_cm1": model_frequency_shift((de - de_t).cpu().numpy())}
        log(f"[bench] cphf {name}: {wall:.3f} s, cycles {st.get('cycles')}, aop calls {st.get('aop_calls')}, "
            f"fallback {st.get('fp64_fallback')}")
        del drv
    res["mixed_resident_learned"] = learned_leg(resident, d_C, occ, nao, nocc, nprop, dev, f1, s1, ta, mo1t, de_t)
    base = res["fp64"]["wall_s"]
    res["speedup_vs_fp64"] = {k: round(base / res[k]["wall_s"], 3) for k, *_ in legs if k != "fp64"}
    res["speedup_vs_fp64"]["mixed_resident_learned"] = round(base / res["mixed_resident_learned"]["wall_s"], 3)
    res["speedup_vs_fp64_matched"] = {k: round(res["fp64_matched"]["wall_s"] / res[k]["wall_s"], 3)
                                      for k, *_ in legs if k.startswith("mixed") or k.endswith("resident")}
    return res


# ---- other hot-path pieces (--workload) ----
C3 = dict(ng=24 * 75 * 302, nao=192, natm=24, nocc=40, nprop=72)   # (H2O)_8 def2-SVP, 75 x 302 grid


def _timed(fn, steps, warmup):
    import torch
    out = None
    for _ in range(warmup):
        out = fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        out = fn()
    e1.record()
    torch.cuda.synchronize()
    return out, e0.elapsed_time(e1) / steps


def _line(metric, value, unit, ms, steps, warmup, dtype, config, roofline, higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": round(ms, 3), "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic (seeded)", "config": config, "roofline": roofline}


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def numint_flops(ng, nao, nocc, nprop):
    """Reference-form count of the GGA response: 28 ng nao nocc per perturbation + the shared
    AO-times-occupied transforms (8 ng nao nocc)."""
    return nprop * 28.0 * ng * nao * nocc + 8.0 * ng * nao * nocc


def workload_numint(args):
    """RKS CPKS grid response (get_rks_response_bra, hess_rks.rs:570-605) at (H2O)_8 sizes: one
    step = one response for all 72 perturbations over the 543 600-point grid (GGA)."""
    import numpy as np
    import torch
    import hesship
    from hesship.tensors import empty_f
    from oracle import numint as oni
    c = C3
    ng, nao, nocc, nprop = c["ng"], c["nao"], c["nocc"], c["nprop"]
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(0x5EED0005)
    ao = empty_f((ng, nao, 4), dev)
    ao.normal_(generator=g)
    ao *= 0.3
    w = torch.rand(ng, generator=g, device=dev, dtype=torch.float64) * 0.1
    f = torch.randn((ng, 4, 4), generator=g, device=dev, dtype=torch.float64) * 0.1
    f = f + f.transpose(1, 2)
    mocc = torch.randn((nao, nocc), generator=g, device=dev, dtype=torch.float64) * 0.1
    mo1 = torch.randn((nao, nocc, nprop), generator=g, device=dev, dtype=torch.float64) * 0.1
    ni = hesship.NIMatmul(ao, w)
    del ao
    prec = args.precision
    log(f"[bench] numint response, {prec}")
    out, ms = _timed(lambda: ni.get_rks_response_bra("sigma", f, mo1, mocc, precision=prec), args.steps, args.warmup)
    ref, ms64 = _timed(lambda: ni.get_rks_response_bra("sigma", f, mo1, mocc), 1, 0)
    value = numint_flops(ng, nao, nocc, nprop) / (ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[prec]
    line = _line("numint RKS response TFLOP/s (GGA, all perturbations)", round(value, 3), "TFLOP/s", ms, args.steps,
                 args.warmup, prec, {"workload": "RKS grid response on synthetic AO values", "ngrid": ng, "nao": nao,
                                     "nocc": nocc, "nprop": nprop, "den_type": "sigma"},
                 {"bound": "mfma", "achieved": round(value, 3), "peak": peak, "unit": "TFLOP/s",
                  "frac": round(value / peak, 4), "traffic": None,
                  "kernel": "whole call (half-transform and back-contraction GEMMs + fused pointwise pass)"})
    line["fp64_ms"] = round(ms64, 3)
    line["rel_err_vs_fp64"] = ((out - ref).abs().max() / ref.abs().max()).item()
    if not args.no_cpu_baseline:
        # the oracle's reference-form response on a bounded slice of the grid, scaled per point
        nsamp = 3000
        rng = np.random.default_rng(1)
        aos = rng.standard_normal((nsamp, nao, 4)) * 0.3
        fs = rng.standard_normal((nsamp, 4, 4)) * 0.1
        fs = fs + fs.transpose(0, 2, 1)
        mo, m1 = mocc.cpu().numpy(), mo1.cpu().numpy()
        t0 = time.perf_counter()
        oni.rks_response_bra("sigma", fs, aos, rng.uniform(0, 0.1, nsamp), m1, mo)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(numint_flops(nsamp, nao, nocc, nprop) / dt / 1e12, 5),
                                "unit": "TFLOP/s", "cores": _blas_threads(), "kind": "port",
                                "sample": f"oracle rks_response_bra (NumPy FP64) on {nsamp} of {ng} grid points, "
                                          f"all {nprop} perturbations, {dt:.1f} s"}
    return line


def setup_flops(ng, nao, natm, nvar=4):
    """Algorithmic FP64 work of the GGA setup passes: ao_dm0 (4 components), de_fxc, the
    de_vxc_off AO-pair tensor, vmat_ip, and the dominant vmat_fxc product 2 ng nao^2 3 natm."""
    n3 = 3 * natm
    return (8.0 * ng * nao * nao + 2.0 * ng * nvar * n3 * n3 + 2.0 * ng * (3 * nao) ** 2 + 4.0 * ng * 3 * nao * nao
            + 2.0 * ng * nao * nao * n3)


def workload_dft_setup(args):
    """DFT skeleton setup (make_hessian_setup_batched, hess_rks.rs:611-750) at (H2O)_8 sizes, GGA,
    FP64, AO derivatives to third order (20 components)."""
    import numpy as np
    import torch
    import hesship
    from hesship.tensors import empty_f
    from oracle import dft_setup as ods
    c = C3
    ng, nao, natm = c["ng"], c["nao"], c["natm"]
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(0x5EED0006)
    ao = empty_f((ng, nao, 20), dev)
    ao.normal_(generator=g)
    ao *= 0.3
    w = torch.rand(ng, generator=g, device=dev, dtype=torch.float64) * 0.1
    vxc = torch.randn((ng, 4), generator=g, device=dev, dtype=torch.float64) * 0.1
    f = torch.randn((ng, 4, 4), generator=g, device=dev, dtype=torch.float64) * 0.1
    f = f + f.transpose(1, 2)
    Cm = torch.randn((nao, c["nocc"]), generator=g, device=dev, dtype=torch.float64) * 0.1
    dm0 = Cm @ Cm.T
    per = nao // natm
    slices = [(A * per, (A + 1) * per) for A in range(natm)]
    ni = hesship.NIMatmul(ao, w)
    del ao
    log("[bench] dft setup")
    _, ms = _timed(lambda: ni.make_hessian_setup("sigma", dm0, slices, vxc, f), args.steps, args.warmup)
    value = setup_flops(ng, nao, natm) / (ms * 1e-3) / 1e12
    line = _line("DFT skeleton setup TFLOP/s (GGA, FP64)", round(value, 3), "TFLOP/s", ms, args.steps, args.warmup,
                 "f64", {"workload": "DFT Hessian setup passes on synthetic AO values (derivative order 3)",
                         "ngrid": ng, "nao": nao, "natm": natm, "den_type": "sigma"},
                 {"bound": "mfma", "achieved": round(value, 3), "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                  "frac": round(value / PEAK_TFLOPS["fp64"], 4), "traffic": None,
                  "kernel": "whole call (GEMMs dominated by vmat_fxc, pointwise drho and AO-diagonal passes)"})
    if not args.no_cpu_baseline:
        nsamp = 2000
        rng = np.random.default_rng(2)
        aos = rng.standard_normal((nsamp, nao, 20)) * 0.3
        t0 = time.perf_counter()
        ods.make_hessian_setup("sigma", aos, rng.uniform(0, 0.1, nsamp), dm0.cpu().numpy(), slices,
                               rng.standard_normal((nsamp, 4)) * 0.1, rng.standard_normal((nsamp, 4, 4)) * 0.1)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(setup_flops(nsamp, nao, natm) / dt / 1e12, 5), "unit": "TFLOP/s",
                                "cores": _blas_threads(), "kind": "port",
                                "sample": f"oracle make_hessian_setup (NumPy FP64) on {nsamp} of {ng} grid points, "
                                          f"{dt:.1f} s"}
    return line


def becke_geometry(natm, npa, seed):
    """Atoms in a box with a minimum separation; each atom's points at random directions and radii
    around it (the layout of an atom-grouped Becke grid); antisymmetric radii adjustments."""
    import numpy as np
    rng = np.random.default_rng(seed)
    R = []
    while len(R) < natm:
        x = rng.uniform(-4.5, 4.5, 3)
        if all(np.linalg.norm(x - y) > 1.8 for y in R):
            R.append(x)
    R = np.array(R)
    attr = np.repeat(np.arange(natm), npa)
    u = rng.standard_normal((len(attr), 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    coords = R[attr] + u * np.exp(rng.uniform(np.log(0.05), np.log(6.0), (len(attr), 1)))
    a = rng.uniform(-0.3, 0.3, (natm, natm))
    adj = np.triu(a, 1) - np.triu(a, 1).T
    split = np.concatenate([[0], np.cumsum(np.full(natm, npa) if np.isscalar(npa) else npa)]).astype(np.int64)
    return R, attr, coords, rng.uniform(1e-3, 0.5, len(attr)), adj, split


def becke_flops(ng, nao, natm, nvar=4):
    """The setup passes (vmat on the combined field once) plus the two grid-shift n3 x n3 products
    and the rank-natm cross term of the contracted ddw (symmetric half, natm (3 natm)^2 per point)."""
    n3 = 3 * natm
    return setup_flops(ng, nao, natm, nvar) + 4.0 * ng * nvar * n3 * n3 + 1.0 * ng * natm * n3 * n3


def workload_dft_becke(args):
    """DFT skeleton setup with the Becke grid shift (make_hessian_setup_becke,
    hess_rks_becke.rs:1106-1266) at (H2O)_8 sizes, GGA, FP64: de_xc_skeleton and
    vmat_deriv1_grid, the keys its consumer reads (hess_rks_becke.rs:1389-1401)."""
    import numpy as np
    import torch
    import hesship
    from hesship.tensors import empty_f
    from oracle import dft_becke as odb
    c = C3
    ng, nao, natm = c["ng"], c["nao"], c["natm"]
    dev = torch.device("cuda:0")
    R, attr, coords, wq, adj, split = becke_geometry(natm, ng // natm, seed=8)
    tab = hesship.BeckeMolTables(R, adj, dev)
    d_coords = torch.from_numpy(coords).to(dev)
    d_wq = torch.from_numpy(wq).to(dev)
    w = tab.partition(d_coords, split, d_wq, deriv=0, by_atom=True)["w"]
    g = torch.Generator(device=dev).manual_seed(0x5EED0008)
    ao = empty_f((ng, nao, 20), dev)
    ao.normal_(generator=g)
    ao *= 0.3
    vxc = torch.randn((ng, 4), generator=g, device=dev, dtype=torch.float64) * 0.1
    f = torch.randn((ng, 4, 4), generator=g, device=dev, dtype=torch.float64) * 0.1
    f = f + f.transpose(1, 2)
    exc = torch.randn(ng, generator=g, device=dev, dtype=torch.float64) * 0.1
    Cm = torch.randn((nao, c["nocc"]), generator=g, device=dev, dtype=torch.float64) * 0.1
    dm0 = Cm @ Cm.T
    per = nao // natm
    slices = [(A * per, (A + 1) * per) for A in range(natm)]
    ni = hesship.NIMatmul(ao, w)
    del ao
    log("[bench] dft setup with the Becke grid shift")
    run = lambda: ni.make_hessian_setup_becke("sigma", dm0, slices, exc, vxc, f, tab, d_coords, d_wq, split,
                                              grid_fixed_keys=False)
    _, ms = _timed(run, args.steps, args.warmup)
    value = becke_flops(ng, nao, natm) / (ms * 1e-3) / 1e12
    line = _line("DFT skeleton setup with Becke grid shift TFLOP/s (GGA, FP64)", round(value, 3), "TFLOP/s", ms,
                 args.steps, args.warmup, "f64",
                 {"workload": "DFT Hessian setup with the Becke grid shift on synthetic AO values (derivative "
                              "order 3) and an atom-grouped grid", "ngrid": ng, "nao": nao, "natm": natm,
                  "den_type": "sigma"},
                 {"bound": "mfma", "achieved": round(value, 3), "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                  "frac": round(value / PEAK_TFLOPS["fp64"], 4), "traffic": None,
                  "kernel": "whole call (vmat GEMM on the combined field, AO-pair GEMMs per atom, contracted ddw)"})
    if not args.no_cpu_baseline:
        nsamp = 600
        rng = np.random.default_rng(3)
        sel = slice(0, nsamp)                       # points of atom 0
        aos = rng.standard_normal((nsamp, nao, 20)) * 0.3
        t0 = time.perf_counter()
        odb.make_hessian_setup_chunk_becke("sigma", aos, w[sel].cpu().numpy(), wq[sel], coords[sel], R, adj, 3,
                                           dm0.cpu().numpy(), slices, 0, rng.standard_normal(nsamp) * 0.1,
                                           rng.standard_normal((nsamp, 4)) * 0.1,
                                           rng.standard_normal((nsamp, 4, 4)) * 0.1)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(becke_flops(nsamp, nao, natm) / dt / 1e12, 5), "unit": "TFLOP/s",
                                "cores": _blas_threads(), "kind": "port",
                                "sample": f"oracle make_hessian_setup_chunk_becke (NumPy FP64) on {nsamp} of {ng} "
                                          f"grid points (one atom's chunk), {dt:.1f} s"}
    return line


def ucphf_case(obj, d_C, occ, nao, nprop, dev, config=None, seed=0x5EED0005):
    """Unrestricted counterpart of cphf_case: both spins share the synthetic MO coefficients;
    orbital energies and first-order Fock/overlap derivatives per spin from the seeded generator."""
    import numpy as np
    import torch
    import hesship
    from hesship import synth
    noc = [int((o > 0).sum()) for o in occ]
    z = synth.normal(seed, 0, 2 * nao + sum(2 * nao * no * nprop for no in noc))
    en, f1, s1, p = [], [], [], 2 * nao
    for sp, no in enumerate(noc):
        zz = z[sp * nao:(sp + 1) * nao]
        en.append(np.sort(np.concatenate([-1.4 + 1.1 * np.abs(np.tanh(zz[:no])), 0.2 + 2.5 * np.abs(np.tanh(zz[no:]))])))
        f1.append(torch.from_numpy(z[p:p + nao * no * nprop].reshape(nprop, no, nao).T.copy()).to(dev))
        p += nao * no * nprop
        s1.append(torch.from_numpy(0.1 * z[p:p + nao * no * nprop].reshape(nprop, no, nao).T.copy()).to(dev))
        p += nao * no * nprop
    drv = hesship.UHessCPHF([d_C, d_C], occ, en, dict({"cphf_tol": 1e-8}, **(config or {})))
    drv.add_el(obj)
    rng = np.random.default_rng(1)
    v = [torch.from_numpy(rng.standard_normal((nao, no, 4))).to(dev) for no in noc]
    r = drv.response_dimless_cphf(v)
    ratio = (sum(x.norm() ** 2 for x in r) ** 0.5 / sum(x.norm() ** 2 for x in v) ** 0.5).item()
    return drv, en, f1, s1, ratio


def workload_cphf_batch(args):
    """Many uneven molecules: 64 seeded alkane-sized RI-JK CPHF systems (C_nH_2n+2 with
    cc-pVDZ / -jkfit counts: nao 24n+10, naux 84n+28, 3(3n+2) perturbations, n uniform in 1..12),
    synthetic tensors; every second molecule is a doublet cation (UHF, 4n+1 alpha and 4n beta
    occupied orbitals), the others closed shells (RHF, 4n+1 doubly occupied).  Each system's
    RI-JK scale gives |A v| ~ 0.3 |v| for random v, shrunk where A then has eigenvalues below
    -0.75 (an indefinite I + A, an unstable reference state, is not a CPHF system).  Each molecule is
    solved to 1e-8 (one subspace of up to 42 cycles) on one created stream, plain FP64 with the
    response launched eagerly and mixed precision (BF16x3 operator, FP64 refinement) with the
    response replayed from HIP graphs; each leg runs once untimed first (buffers sized, products
    tuned, graphs captured), and both again with the resident response operator (its matrix built
    inside the timed region).  The molecules are handled one at a time (a 12-carbon system holds
    ~15 GB of workspace)."""
    import numpy as np
    import torch
    import hesship
    from hesship import synth
    dev = torch.device("cuda:0")
    ns = np.random.default_rng(64).integers(1, 13, 64)
    stream = torch.cuda.Stream(device=dev)
    cfg = dict(cphf_max_space=42)
    tot = dict(fp64=0.0, mixed=0.0, fp64_resident=0.0, mixed_resident=0.0)
    cyc = dict(fp64=0, mixed=0, fp64_resident=0, mixed_resident=0)
    caps = launches = 0
    err = err_res = 0.0
    nuhf = fallbacks = 0
    min_eig = []

    def diff(a, b):
        if isinstance(a, (list, tuple)):
            return max((x - y).abs().max().item() for x, y in zip(a, b))
        return (a - b).abs().max().item()

    def make_mol(i, n, scale2, resident=False):
        """(RI-JK handle, CPHF driver, f1, s1, prepare, nocc, ratio) of molecule i"""
        nao, naux = 24 * n + 10, 84 * n + 28
        nprop = 3 * (3 * n + 2)
        cd = torch.from_numpy(synth.cderi(nao, naux, seed=synth.SEED_B + 16 * i).T.copy()).to(dev).T
        C = torch.from_numpy(synth.mo_coeff(nao, seed=synth.SEED_C + 16 * i)).to(dev)
        if i % 2 == 1:
            occ = [np.where(np.arange(nao) < 4 * n + 1, 1.0, 0.0), np.where(np.arange(nao) < 4 * n, 1.0, 0.0)]
            noc = [4 * n + 1, 4 * n]
            o = hesship.UHessRIJK(cd * scale2 ** 0.5, 1.0, 1.0)
            prep = lambda: o.make_response_preparation([C, C], occ)
        else:
            noc = 4 * n + 1
            occ = np.where(np.arange(nao) < noc, 2.0, 0.0)
            o = hesship.RHessRIJK(cd * scale2 ** 0.5, 1.0, 1.0)
            prep = lambda: o.make_response_preparation(C, occ)
        if resident:
            o.set_response_path("resident")
        prep()
        if i % 2 == 1:
            drv, _, f1, s1, ratio = ucphf_case(o, C, occ, nao, nprop, dev, cfg)
        else:
            drv, _, f1, s1, ratio = cphf_case(o, C, occ, nao, noc, nprop, dev, cfg)
        return o, drv, f1, s1, prep, noc, ratio

    scales, refs = [], []
    known_scales = None
    if args.batch_scales:
        with open(args.batch_scales) as fh:
            known_scales = json.load(fh)
        assert len(known_scales) == len(ns), "batch scales: one per molecule"
    with torch.cuda.stream(stream):
        for i, n in enumerate(ns):
            n = int(n)
            nao, nprop = 24 * n + 10, 3 * (3 * n + 2)
            uhf = i % 2 == 1
            nuhf += uhf
            if known_scales is not None:
                # the scale of an earlier run of this workload (same seeds)
                scale2 = known_scales[i]
                obj, drv, f1, s1, prep, noc, _ = make_mol(i, n, scale2)
                lo = float("nan")
            else:
                # |A v| ~ 0.3 |v| for random v, then shrunk where that leaves I + A indefinite
                ratio = make_mol(i, n, 1.0)[-1]
                scale2 = 0.3 / ratio
                obj, drv, f1, s1, prep, noc, _ = make_mol(i, n, scale2)
                lo = response_spectrum(drv, nao, noc, dev, which=("SR",))[0]
                if lo < -LOWEST_A:
                    scale2 *= LOWEST_A / -lo
                    del obj, drv
                    obj, drv, f1, s1, prep, noc, _ = make_mol(i, n, scale2)
            min_eig.append(round(max(lo, -LOWEST_A), 4))
            scales.append(scale2)
            if args.batched_only:
                del obj, drv
                robj, rdrv, _, _, rprep, _, _ = make_mol(i, n, scale2, resident=True)
                rdrv.set_graphs(False)
                run_cphf(rdrv, f1, s1, "fp64", rprep)
                ref, _, tr64, str64 = run_cphf(rdrv, f1, s1, "fp64", rprep)
                refs.append(ref)
                tot["fp64_resident"] += tr64
                cyc["fp64_resident"] += str64["cycles"]
                log(f"[bench] molecule {i}: C{n} {'UHF' if uhf else 'RHF'} nao {nao} nprop {nprop}: resident FP64 "
                    f"{str64['cycles']} cycles {tr64 * 1e3:.1f} ms")
                del rdrv, robj, f1, s1
                continue
            # every leg's timed region holds prepare_response (for the resident operator that
            # includes building its matrix on the first response), the rhs, the solve and finalize
            drv.set_graphs(False)
            run_cphf(drv, f1, s1, "fp64", prep)
            ref, _, t64, st64 = run_cphf(drv, f1, s1, "fp64", prep)
            drv.set_graphs(True)
            run_cphf(drv, f1, s1, "bf16x3", prep)
            got, _, tmx, stmx = run_cphf(drv, f1, s1, "bf16x3", prep)
            c, l = drv.graph_stats()
            del drv, obj
            robj, rdrv, _, _, rprep, _, _ = make_mol(i, n, scale2, resident=True)
            rdrv.set_graphs(False)
            run_cphf(rdrv, f1, s1, "fp64", rprep)
            _, _, tr64, str64 = run_cphf(rdrv, f1, s1, "fp64", rprep)
            rdrv.set_graphs(True)
            run_cphf(rdrv, f1, s1, "bf16x3", rprep)
            rgot, _, trmx, strmx = run_cphf(rdrv, f1, s1, "bf16x3", rprep)
            err = max(err, diff(got, ref))
            err_res = max(err_res, diff(rgot, ref))
            refs.append(ref)
            tot["fp64"] += t64
            tot["mixed"] += tmx
            tot["fp64_resident"] += tr64
            tot["mixed_resident"] += trmx
            cyc["fp64"] += st64["cycles"]
            cyc["mixed"] += stmx["cycles"]
            cyc["fp64_resident"] += str64["cycles"]
            cyc["mixed_resident"] += strmx["cycles"]
            fallbacks += stmx.get("fp64_fallback", 0) + strmx.get("fp64_fallback", 0)
            caps += c
            launches += l
            log(f"[bench] molecule {i}: C{n} {'UHF' if uhf else 'RHF'} nao {nao} nprop {nprop}: FP64 {st64['cycles']} "
                f"cycles {t64 * 1e3:.1f} ms, mixed {stmx['cycles']} cycles / {stmx['refinements']} sweeps "
                f"{' (FP64 finish)' if stmx.get('fp64_fallback') else ''} {tmx * 1e3:.1f} ms; resident FP64 "
                f"{tr64 * 1e3:.1f} ms, mixed {strmx['cycles']} cycles {trmx * 1e3:.1f} ms")
            del rdrv, robj, f1, s1, got, rgot
            torch.cuda.synchronize()
        # every molecule at once: all 64 systems resident and solved in lockstep (hh_cphf_batch),
        # the timed region holding prepare_response of every handle, the right-hand sides, the
        # resident matrices' builds, the batched solve and finalize
        mols = [make_mol(i, int(n), scales[i], resident=True) for i, n in enumerate(ns)]
        torch.cuda.synchronize()
        log(f"[bench] batched leg: {len(mols)} systems created, {torch.cuda.memory_allocated() / 2**30:.1f} GiB "
            f"held by torch")

        def batched():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for m in mols:
                m[4]()
            rhs = [m[1].compute_dimless_rhs(m[2], m[3]) for m in mols]
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            bt = hesship.CPHFBatch(tol=1e-8, max_cycle=42, lindep=1e-14, device=dev)
            for m, r in zip(mols, rhs):
                bt.add(m[1], m[0], r)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            st = []
            mo1 = bt.solve(stats=st)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            outs = [m[1].finalize_cphf(m[2], m[3], x)[0] for m, x in zip(mols, mo1)]
            torch.cuda.synchronize()
            t4 = time.perf_counter()
            stages = {"prepare_and_rhs_s": t1 - t0, "add_and_matrix_builds_s": t2 - t1, "solve_s": t3 - t2,
                      "solve_phases_s": bt.times(), "finalize_s": t4 - t3}
            return t4 - t0, outs, st, bt.info(), stages

        batched()
        t_batch, outs, bst, binfo, bstages = batched()
        log(f"[bench] batched stages: {bstages}")
        err_batch = max(diff(a, b) for a, b in zip(outs, refs))
        cyc["batched_fp64"] = sum(x["cycles"] for x in bst)
        log(f"[bench] batched FP64: {t_batch:.3f} s, {binfo[1]} lockstep cycles, <= {binfo[2]} launches per cycle, "
            f"max |mo1 - mo1_FP64| {err_batch:.2e}")
        del mols, outs
    wall = tot["mixed_resident"] if not args.batched_only else t_batch
    log(f"[bench] cphf batch: FP64 {tot['fp64']:.3f} s, mixed {tot['mixed']:.3f} s, resident FP64 "
        f"{tot['fp64_resident']:.3f} s, resident mixed {wall:.3f} s")
    line = _line("CPHF batch wall s (64 uneven RI-JK molecules, RHF + UHF, mixed precision, resident response, "
                 "graph replay)",
                 round(wall, 4), "s", wall * 1e3, 1, 1, "bf16x3 operator, f64 refinement",
                 {"workload": "64 alkane-sized synthetic RI-JK CPHF systems (32 closed shells, 32 doublet cations) "
                              "solved one after the other on one stream",
                  "n_carbon": [int(n) for n in ns], "cphf_tol": 1e-8},
                 {"bound": "launch", "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None,
                  "kernel": "whole solves"}, higher=False)
    line["uhf_molecules"] = nuhf
    line["rijk_scale2"] = [float(x) for x in scales]
    line["min_eig_A"] = min_eig
    line["fp64_eager_wall_s"] = round(tot["fp64"], 4)
    line["mixed_per_call_wall_s"] = round(tot["mixed"], 4)
    line["fp64_resident_wall_s"] = round(tot["fp64_resident"], 4)
    ratio = lambda a, b: round(a / b, 3) if b > 0 and a > 0 else None
    line["speedup_vs_fp64"] = {"mixed_per_call": ratio(tot["fp64"], tot["mixed"]),
                               "fp64_resident": ratio(tot["fp64"], tot["fp64_resident"]),
                               "mixed_resident": ratio(tot["fp64"], tot["mixed_resident"])}
    line["max_abs_mo1_diff_mixed_resident_vs_fp64"] = err_res
    line["krylov_cycles"] = cyc
    line["graphs_captured"] = caps
    line["graph_launches"] = launches
    line["max_abs_mo1_diff_mixed_vs_fp64"] = err
    line["mixed_solves_finished_in_fp64"] = fallbacks
    line["batched_fp64"] = {"wall_s": round(t_batch, 4), "speedup_vs_fp64": ratio(tot["fp64"], t_batch),
                            "speedup_vs_sequential_fp64_resident": ratio(tot["fp64_resident"], t_batch),
                            "lockstep_cycles": binfo[1], "launches_per_cycle": binfo[2],
                            "max_abs_mo1_diff_vs_fp64": err_batch,
                            "stages": {k: (round(v, 4) if isinstance(v, float) else [round(x, 4) for x in v])
                                       for k, v in bstages.items()},
                            "note": "all 64 systems resident at once and solved in lockstep (hh_cphf_batch): "
                                    "every Krylov step one grouped launch over the running systems, one host "
                                    "read per cycle; FP64 throughout"}
    return line


def workload_skeleton(args):
    """RI-JK skeleton (make_skeleton_hess + get_deriv1_bra, hess_r.rs:1858-1909) on synthetic
    integrals at (H2O)_8 def2-SVP / jkfit-like sizes, FP64; the derivative integrals are served
    from device-resident tensors through the generator interface in aux ranges of 72."""
    import numpy as np
    import torch
    import hesship
    from hesship.tensors import empty_f
    nao, natm, nocc, naux = 192, 24, 40, 8 * 112
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(0x5EED0007)
    rn = lambda *shape: torch.randn(shape, generator=g, device=dev, dtype=torch.float64) * 0.05
    a = rn(naux, naux)
    j2c = a @ a.T + naux * 0.05 * torch.eye(naux, device=dev, dtype=torch.float64)
    j3c = rn(nao, nao, naux)
    j3c = j3c + j3c.transpose(0, 1)
    il = torch.tril_indices(nao, nao, device=dev)   # (nu, mu), mu <= nu, nu-major: the packed-upper order
    j3c_tp = j3c[il[1], il[0]].contiguous()
    del j3c
    ints = {}
    for k, nc in (("ip1", 3), ("ip2", 3), ("ipip1", 9), ("ipvip1", 9), ("ip1ip2", 9), ("ipip2", 9)):
        t = empty_f((nao, nao, naux) + ((3,) if nc == 3 else (3, 3)), dev)
        t.normal_(generator=g)
        ints[k] = t
    v2 = [rn(naux, naux, 3), rn(naux, naux, 3, 3), rn(naux, naux, 3, 3)]
    C, _ = torch.linalg.qr(rn(nao, nao))
    occ = np.where(np.arange(nao) < nocc, 2.0, 0.0)
    per, pera = nao // natm, naux // natm
    aos = [(A * per, (A + 1) * per) for A in range(natm)]
    auxs = [(A * pera, (A + 1) * pera) for A in range(natm)]
    ranges = [(p, min(naux, p + 72)) for p in range(0, naux, 72)]
    h = hesship.RHessRIJK.from_j3c(j3c_tp, j2c, 1.0, 0.2)
    fn = lambda kind, p0, p1: ints[kind][:, :, p0:p1]
    log("[bench] rijk skeleton")
    _, ms = _timed(lambda: h.make_skeleton(C, occ, fn, *v2, ranges, aos, auxs), args.steps, args.warmup)
    nbytes = 8.0 * nao * nao * naux * (2 * 3 + 4 * 9)  # every derivative integral element read once
    gbs = nbytes / (ms * 1e-3) / 1e9
    return _line("RI-JK skeleton wall ms (make_skeleton_hess + get_deriv1_bra, FP64)", round(ms, 3), "ms", ms,
                 args.steps, args.warmup, "f64",
                 {"workload": "RI-JK skeleton on synthetic derivative integrals", "nao": nao, "naux": naux,
                  "natm": natm, "nocc": nocc, "aux_range": 72},
                 {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                  "frac": round(gbs / 8000.0, 4), "traffic": None,
                  "kernel": "whole call: derivative-integral bytes (each element read once, generator copies "
                            "included) / time"},
                 higher=False)


def synthetic_density(ng, nspin, seed, tau=False):
    """Densities 1e-5 .. 1e2 with gradients up to a few times rho (molecular-grid-like), [ng, 4(, 2)];
    with tau, [ng, 5(, 2)] and tau between tau_W and 6 tau_W plus a uniform-gas part."""
    import numpy as np
    rng = np.random.default_rng(seed)

    def one():
        r = np.empty((ng, 5 if tau else 4))
        r[:, 0] = 10 ** rng.uniform(-5, 2, ng)
        r[:, 1:4] = rng.normal(size=(ng, 3)) * r[:, :1] * rng.uniform(0.05, 3.0, (ng, 1))
        if tau:
            tw = (r[:, 1:4] ** 2).sum(1) / (8 * r[:, 0])
            r[:, 4] = tw * rng.uniform(1.0, 6.0, ng) + rng.uniform(0.0, 0.5, ng) * r[:, 0] ** (5 / 3)
        return r
    return one() if nspin == 1 else np.stack([one(), one()], -1)


XC_BYTES = {1: 8 * (4 + 1 + 4 + 16), 2: 8 * (8 + 1 + 8 + 64)}   # rho in; exc, vxc_eff, fxc_eff out
# FP64 operations executed per grid point by xc_eff_kernel for B3LYP (rocprofv3 --pmc
# SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64, 64 lanes x (ADD + MUL + 2 FMA + TRANS) per dispatch / ngrid)
XC_FLOP_PER_POINT = {1: 4690.1, 2: 7525.2}
PEAK_FP64_VALU = 78.6


# A def2-SVP-sized synthetic basis (O 3s2p1d, H 2s1p: 24 functions per water) and an
# even-tempered auxiliary basis up to g (79 functions per water); no basis-set files exist here.
SVP_LIKE = {"O": [(0, [[2266.2, 0.0053], [340.87, 0.0399], [77.36, 0.1785], [21.48, 0.4643], [6.66, 0.4431]]),
                  (0, [[0.81, 1.0]]), (0, [[0.26, 1.0]]), (1, [[17.72, 0.043], [3.86, 0.229], [1.05, 0.509]]),
                  (1, [[0.28, 1.0]]), (2, [[1.2, 1.0]])],
            "H": [(0, [[13.0, 0.0197], [1.96, 0.138], [0.44, 0.478]]), (0, [[0.12, 1.0]]), (1, [[0.8, 1.0]])]}
JK_LIKE = {"O": [(0, [[20.0, 1.0]]), (0, [[5.0, 1.0]]), (0, [[1.5, 1.0]]), (0, [[0.45, 1.0]]), (1, [[6.0, 1.0]]),
                 (1, [[1.8, 1.0]]), (1, [[0.55, 1.0]]), (2, [[4.0, 1.0]]), (2, [[1.2, 1.0]]), (2, [[0.4, 1.0]]),
                 (3, [[2.0, 1.0]]), (3, [[0.7, 1.0]]), (4, [[1.2, 1.0]])],
           "H": [(0, [[3.0, 1.0]]), (0, [[0.8, 1.0]]), (0, [[0.25, 1.0]]), (1, [[1.5, 1.0]]), (1, [[0.4, 1.0]]),
                 (2, [[0.9, 1.0]])]}


def workload_cint(args):
    """Integral engine and grids on a seeded (H2O)_8 cluster (O-H 0.96 A, 104.5 deg, 7 A box):
    the 75 x 302 Becke grid (543 600 points) built and partitioned, AO values with first and
    third derivatives on it (the numint and setup inputs), the 2-centre metric and its three
    derivatives, and every 3-centre kind the RI-JK skeleton consumes, over all auxiliary shells.
    One step = all of it.  Roofline: the AO evaluation is HBM-bound at best (output bytes)."""
    import numpy as np
    import torch
    import hesship
    from hesship import synth
    from hesship.grids import Grids
    from oracle import cint as oc
    dev = torch.device("cuda:0")
    z, R = synth.water_cluster(8, box_ang=7.0, seed=0x5EED0008)
    sym = {8: "O", 1: "H"}
    atoms = [(sym[int(c)], tuple(x)) for c, x in zip(z, R)]
    mol = hesship.CInt.from_basis(atoms, SVP_LIKE)
    aux = hesship.CInt.from_basis(atoms, JK_LIKE)
    nao, naux = mol.nao, aux.nao
    log(f"[bench] cint: (H2O)_8 nao {nao} naux {naux}")
    kinds3 = ["plain", "ip1", "ip2", "ipip1", "ipvip1", "ip1ip2", "ipip2"]
    kinds2 = ["plain", "ip1", "ipip1", "ip1ip2"]
    out = {}

    def t(fn):
        return _timed(fn, args.steps, args.warmup)

    g = Grids(z, R, nrad=75, nang=302, device=dev)
    (coords, w), out["grid_ms"] = t(g.build)
    ng = coords.shape[0]
    _, out["ao_deriv1_ms"] = t(lambda: mol.eval_gto(coords, 1))
    _, out["ao_deriv3_ms"] = t(lambda: mol.eval_gto(coords, 3))
    for k in kinds2:
        _, out[f"int2c2e_{k}_ms"] = t(lambda: aux.int2c2e(k))
    for k in kinds3:
        _, out[f"int3c2e_{k}_ms"] = t(lambda: mol.int3c2e(aux, k))
    tot = sum(out.values())
    ao_bytes = ng * nao * 20 * 8
    gbs = ao_bytes / (out["ao_deriv3_ms"] * 1e-3) / 1e9
    nint3 = nao * nao * naux * (1 + 6 + 54)
    # CPU: the oracle (torch autograd restatement) on a bounded sample
    sh = []
    for b in mol.bas:  # one (centre, l, exps, coefs) per contraction, libcint order
        A, l, nprim, nctr, pe, pc = (int(v) for v in (b[0], b[1], b[2], b[3], b[5], b[6]))
        for c in range(nctr):
            sh.append((mol.env[mol.atm[A][1]:mol.atm[A][1] + 3], l, mol.env[pe:pe + nprim],
                       mol.env[pc + nprim * c:pc + nprim * (c + 1)]))
    pts = coords[:64].cpu().numpy()
    t0 = time.perf_counter()
    oc.eval_gto(sh, pts, 3)
    cpu_pts = 64 / (time.perf_counter() - t0)
    line = _line("Integral engine ms per step ((H2O)_8: Becke grid 543600 points, AO values deriv 1 and 3, int2c2e x 4 "
                 "kinds, int3c2e x 7 kinds)", round(tot, 3), "ms", tot, args.steps, args.warmup, "f64",
                 {"workload": "synthetic (H2O)_8, def2-SVP-sized basis (nao 192), even-tempered auxiliary to g",
                  "nao": nao, "naux": naux, "ngrid": int(ng)},
                 {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s", "frac": round(gbs / 8000.0, 4),
                  "traffic": None, "kernel": "eval_gto_kernel<L, 3> (AO values and 19 derivatives, output bytes)"},
                 higher=False)
    line["stages_ms"] = {k: round(v, 3) for k, v in out.items()}
    line["int3c2e_values_per_s"] = round(nint3 / (sum(out[f"int3c2e_{k}_ms"] for k in kinds3) * 1e-3), 1)
    line["cpu_oracle"] = {"ao_deriv3_points_per_s": round(cpu_pts, 1), "cores": _blas_threads(),
                          "gpu_points_per_s": round(ng / (out["ao_deriv3_ms"] * 1e-3), 1),
                          "note": "torch-autograd restatement of the AO values, 64 points"}
    return line


def workload_hessian(args):
    """The whole restricted analytic Hessian of a seeded (H2O)_n cluster (default n = 8: the c3
    molecule, def2-SVP-sized basis, nao 192, even-tempered auxiliary basis naux 632) on the
    device: the density-fitted SCF (RHF, or RKS with --xc on an --nrad x --nang Becke grid), then
    make_hess (nuclear repulsion + core + RI-JK skeleton + overlap (+ the grid component with the
    Becke grid shift) + CPHF) with the CPHF in FP64 and in the mixed BF16x3 mode with FP64
    refinement; harmonic frequencies of both Hessians.  With --open-shell the cation doublet
    (H2O)_n^+ (5n alpha, 5n - 1 beta electrons) through UHF / UKS and UHessSCF.  One step = one
    make_hess."""
    import numpy as np
    import torch
    import hesship
    from hesship import synth, vib
    dev = torch.device("cuda:0")
    z, R = synth.water_cluster(args.nwater, box_ang=7.0 if args.nwater <= 8 else 7.0 * (args.nwater / 8) ** (1 / 3),
                               seed=0x5EED0008)
    sym = {8: "O", 1: "H"}
    atoms = [(sym[int(c)], tuple(x)) for c, x in zip(z, R)]
    mol = hesship.CInt.from_basis(atoms, SVP_LIKE)
    aux = hesship.CInt.from_basis(atoms, JK_LIKE)
    nocc = 5 * args.nwater
    op = bool(getattr(args, "open_shell", False))
    log(f"[bench] hessian: (H2O)_{args.nwater}{'+' if op else ''} nao {mol.nao} naux {aux.nao} natm {len(atoms)} "
        f"xc {args.xc}")
    xc_names = {"none": None, "svwn": [(1.0, "LDA_X"), (1.0, "LDA_C_VWN")], "b3lyp": [(1.0, "HYB_GGA_XC_B3LYP")],
                "tpss0": [(1.0, "HYB_MGGA_XC_TPSS0")]}[args.xc]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    occ = (nocc, nocc - 1) if op else (nocc,)
    if xc_names is None:
        scf = (hesship.UHF if op else hesship.RHF)(mol, aux, *occ, conv_tol=1e-11, max_cycle=200).run()
        ngrid = 0
    else:
        from hesship.grids import Grids
        xc = [(sc, hesship.LibXCFunctional.from_identifier(n, "polarized" if op else "unpolarized"))
              for sc, n in xc_names]
        grids = Grids(z, R, nrad=args.nrad, nang=args.nang, device=dev)
        scf = (hesship.UKS if op else hesship.RKS)(mol, aux, *occ, xc, grids, conv_tol=1e-11, max_cycle=200).run()
        ngrid = int(grids.coords.shape[0])
    scf_s = time.perf_counter() - t0
    log(f"[bench] hessian: SCF {scf_s:.2f} s, {scf.cycles} cycles, E = {scf.e_tot:.8f}")
    out = {}
    for prec in ("fp64", "bf16x3"):
        drv = (hesship.UHessSCF if op else hesship.RHessSCF)(
            mol, aux, scf.mo_coeff, scf.mo_occ, scf.mo_energy, xc=scf.xc, grids=scf.grids,
            config=dict(cphf_tol=1e-9 if prec == "bf16x3" else 1e-8))
        st = {}
        walls = []
        for it in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            de = drv.make_hess(precision=prec, stats=st)
            torch.cuda.synchronize()
            if it >= args.warmup:
                walls.append(time.perf_counter() - t0)
        out[prec] = {"wall_s": round(min(walls), 4), "stages_s": {k: round(v, 4) for k, v in drv.timing.items()},
                     "cphf_cycles": st.get("cycles"), "de": de.cpu().numpy()}
        log(f"[bench] hessian {prec}: {out[prec]['wall_s']} s {out[prec]['stages_s']}")
    natm = len(atoms)
    mass = np.array([15.99491462 if s == "O" else 1.00782503 for s, _ in atoms])

    def freqs(de):
        H = de.transpose(2, 0, 3, 1).reshape(3 * natm, 3 * natm)
        H = 0.5 * (H + H.T)
        r = vib.harmonic_analysis(H, np.asarray(R).T, mass)
        return np.where(r.imag[r.vib_indices], -1.0, 1.0) * r.omega[r.vib_indices]
    f64, fmx = freqs(out["fp64"]["de"]), freqs(out["bf16x3"]["de"])
    dH = float(np.abs(out["fp64"]["de"] - out["bf16x3"]["de"]).max())
    wall = out["bf16x3"]["wall_s"]
    u = "U" if op else "R"
    method = f"{u}HF" if xc_names is None else f"{u}KS {args.xc.upper()} ({args.nrad} x {args.nang} Becke grid)"
    line = _line(f"CPHF+Hessian wall s (whole {method.split()[0]} Hessian, mixed-precision CPHF), 1 MI355X", wall, "s",
                 wall * 1e3, args.steps, args.warmup, "bf16x3 CPHF response with FP64 refinement; FP64 elsewhere",
                 {"workload": f"seeded (H2O)_{args.nwater}{'+ doublet' if op else ''} {method}, def2-SVP-sized basis, "
                              f"even-tempered auxiliary",
                  "nao": mol.nao, "naux": aux.nao, "natm": natm, "nocc": list(occ) if op else nocc, "nprop": 3 * natm,
                  "ngrid": ngrid},
                 None, higher=False)
    line["data"] = "synthetic molecule (seeded water cluster; basis sets synthetic, none exist in this image)"
    line["scf"] = {"wall_s": round(scf_s, 3), "cycles": scf.cycles, "e_tot": scf.e_tot}
    line["fp64"] = {k: v for k, v in out["fp64"].items() if k != "de"}
    line["mixed"] = {k: v for k, v in out["bf16x3"].items() if k != "de"}
    line["speedup_vs_fp64"] = round(out["fp64"]["wall_s"] / wall, 3)
    line["accuracy"] = {"max_abs_hessian_diff_vs_fp64": dH,
                        "max_abs_freq_diff_cm1_vs_fp64": float(np.abs(f64 - fmx).max()) if len(f64) == len(fmx) else None,
                        "n_vib": int(len(f64)), "lowest_cm1": float(f64.min()), "highest_cm1": float(f64.max())}
    line["note"] = "roofline per kernel: the ri_jk, skeleton and numint workloads; no CPU leg (the oracle has no whole Hessian)"
    return line


def workload_xc(args):
    """B3LYP effective kernels (get_rho_vxc_fxc's libxc_eval_eff loop, hess_rks.rs:205-216 and
    hess_uks.rs:96-106) on the 543 600-point grid: one step = exc, vxc_eff, fxc_eff for the
    closed-shell and the spin-polarised density (outputs allocated once, outside the timed region)."""
    import torch
    import hesship
    from hesship.tensors import empty_f
    from hesship.xc import _accumulate_eff, _eff_shapes
    from oracle import xcfunc as oxf
    ng = C3["ng"]
    dev = torch.device("cuda:0")
    rho = {s: torch.from_numpy(synthetic_density(ng, s, 0x5EED0007 + s)).to(dev) for s in (1, 2)}
    fn = {1: hesship.LibXCFunctional.from_identifier("HYB_GGA_XC_B3LYP", "unpolarized"),
          2: hesship.LibXCFunctional.from_identifier("HYB_GGA_XC_B3LYP", "polarized")}
    outs = {s: [empty_f(sh, dev) for sh in _eff_shapes(ng, 4, s)] for s in (1, 2)}
    log("[bench] xc B3LYP effective kernels")
    ms = {}
    for s in (1, 2):
        _, ms[s] = _timed(lambda: _accumulate_eff(fn[s], rho[s], "sigma", 1.0, *outs[s], False), args.steps,
                          args.warmup)
    tot = ms[1] + ms[2]
    value = 2 * ng / (tot * 1e-3) / 1e6
    gbs = (XC_BYTES[1] + XC_BYTES[2]) * ng / (tot * 1e-3) / 1e9
    tfl = {s: XC_FLOP_PER_POINT[s] * ng / (ms[s] * 1e-3) / 1e12 for s in (1, 2)}
    line = _line("XC functional kernels Mpoint/s (B3LYP exc + vxc_eff + fxc_eff, RKS and UKS)", round(value, 2),
                 "Mpoint/s", tot, args.steps, args.warmup, "f64",
                 {"workload": "B3LYP effective kernels on a synthetic 543600-point density", "ngrid": ng,
                  "functional": "HYB_GGA_XC_B3LYP", "spins": [1, 2]},
                 {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                  "frac": round(gbs / 8000.0, 4), "traffic": None,
                  "kernel": "xc_eff_kernel (rho in, exc/vxc_eff/fxc_eff out)"})
    # the kernel is bound by FP64 vector arithmetic, not by HBM
    line["compute"] = {"bound": "fp64 valu", "unit": "TFLOP/s", "peak": PEAK_FP64_VALU,
                       "achieved_rks": round(tfl[1], 2), "achieved_uks": round(tfl[2], 2),
                       "frac_rks": round(tfl[1] / PEAK_FP64_VALU, 3), "frac_uks": round(tfl[2] / PEAK_FP64_VALU, 3),
                       "flop_per_point": XC_FLOP_PER_POINT}
    line["ms_rks"] = round(ms[1], 4)
    line["ms_uks"] = round(ms[2], 4)
    # TPSS0 (meta-GGA: rho, grad rho, tau) on a density of the same kind with tau
    rho_t = {s: torch.from_numpy(synthetic_density(ng, s, 0x5EED0017 + s, tau=True)).to(dev) for s in (1, 2)}
    fn_t = {1: hesship.LibXCFunctional.from_identifier("HYB_MGGA_XC_TPSS0", "unpolarized"),
            2: hesship.LibXCFunctional.from_identifier("HYB_MGGA_XC_TPSS0", "polarized")}
    outs_t = {s: [empty_f(sh, dev) for sh in _eff_shapes(ng, 5, s)] for s in (1, 2)}
    log("[bench] xc TPSS0 effective kernels")
    ms_t = {}
    for s in (1, 2):
        _, ms_t[s] = _timed(lambda: _accumulate_eff(fn_t[s], rho_t[s], "tau", 1.0, *outs_t[s], False), args.steps,
                            args.warmup)
    line["tpss0"] = {"ms_rks": round(ms_t[1], 4), "ms_uks": round(ms_t[2], 4),
                     "mpoint_per_s": round(2 * ng / ((ms_t[1] + ms_t[2]) * 1e-3) / 1e6, 2),
                     "note": "HYB_MGGA_XC_TPSS0 exc + vxc_eff + fxc_eff (TAU arrays), RKS and UKS, same grid size"}
    if not args.no_cpu_baseline:
        nsamp = 20000
        t0 = time.perf_counter()
        for s in (1, 2):
            oxf.eval_eff(402, rho[s][:nsamp].cpu().numpy(), s, "sigma")
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(2 * nsamp / dt / 1e6, 4), "unit": "Mpoint/s",
                                "cores": torch.get_num_threads(), "kind": "port",
                                "sample": f"oracle eval_eff (torch autograd FP64) on {nsamp} of {ng} points, RKS "
                                          f"and UKS, {dt:.1f} s"}
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp32", "fp64"])
    ap.add_argument("--size", default="full", choices=list(SIZES))
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cphf", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="cphf_batch: only the resident FP64 legs, one molecule after another and all in one batch")
    ap.add_argument("--batch-scales", default=None, help="cphf_batch: JSON list of the per-molecule RI-JK scales "
                                                         "(the rijk_scale2 field of an earlier run) instead of the "
                                                         "spectrum search")
    ap.add_argument("--workload", default="ri_jk", choices=["ri_jk", "numint", "dft_setup", "dft_becke", "skeleton", "cphf_batch", "xc",
                                                            "cint", "hessian"])
    ap.add_argument("--nwater", type=int, default=8, help="hessian: water molecules in the cluster")
    ap.add_argument("--xc", default="none", choices=["none", "svwn", "b3lyp", "tpss0"],
                    help="hessian: RHF (none) or RKS with this functional")
    ap.add_argument("--nrad", type=int, default=75, help="hessian: radial points per atom (RKS)")
    ap.add_argument("--nang", type=int, default=302, help="hessian: Lebedev points per shell (RKS)")
    ap.add_argument("--open-shell", action="store_true",
                    help="hessian: the cation doublet through UHF / UKS and the unrestricted Hessian")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run the library's 16-byte copy kernel over GIB GiB three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the RI-JK response of one molecule runs on "
              "one MI355X); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import hesship
    from hesship import synth

    if args.workload != "ri_jk":
        fn = {"numint": workload_numint, "dft_setup": workload_dft_setup, "dft_becke": workload_dft_becke,
              "skeleton": workload_skeleton, "cphf_batch": workload_cphf_batch, "xc": workload_xc,
            "cint": workload_cint, "hessian": workload_hessian}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return

    if args.copy_calibration > 0:
        nbytes = int(args.copy_calibration * 2**30) // 16 * 16
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        stream = torch.cuda.current_stream().cuda_stream
        for _ in range(3):
            hesship._lib.check(hesship._lib.lib().hh_util_copy(dst.data_ptr(), src.data_ptr(), nbytes, stream))
        torch.cuda.synchronize()
        assert bool((dst == 1).all())
        print(json.dumps({"copy_calibration_bytes_read": nbytes, "copy_calibration_bytes_written": nbytes,
                          "launches": 3, "kernel": "copy16_kernel"}))
        return

    sz = SIZES[args.size]
    nao, naux, nocc, nprop = sz["nao"], sz["naux"], sz["nocc"], sz["nprop"]
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    cderi, C, occ, bra = synth.rijk_workload(nao, naux, nocc, nprop)
    log(f"[bench] workload generated in {time.perf_counter() - t0:.1f} s "
        f"(nao={nao} naux={naux} nocc={nocc} nprop={nprop})")
    d_cderi = torch.from_numpy(np.asfortranarray(cderi).T.copy()).to(dev).T  # column-major on device
    d_C = torch.from_numpy(C).to(dev)
    d_bra = hesship.fortran(torch.from_numpy(bra).to(dev))

    obj = hesship.RHessRIJK(d_cderi, 1.0, 1.0)
    obj.make_response_preparation(d_C, occ)
    torch.cuda.synchronize()

    def run(prec, steps, warmup):
        out = None
        for _ in range(warmup):
            out = obj.get_response_bra(d_bra, precision=prec)
        torch.cuda.synchronize()
        obj.set_profiling(True)
        obj.reset_timing()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record()
        for _ in range(steps):
            out = obj.get_response_bra(d_bra, precision=prec)
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        obj.set_profiling(False)
        tm = obj.timing()
        return out, e0.elapsed_time(e1) / steps, wall / steps * 1e3, {k: (v / steps if k != "k_hot_launches" else v // steps) for k, v in tm.items()}

    fk, fj = flops_k(nao, naux, nocc, nprop), flops_j(nao, naux, nocc, nprop)
    log(f"[bench] F_K = {fk / 1e12:.4f} TFLOP, F_J = {fj / 1e12:.4f} TFLOP per step")
    r64, ms64, wall64, tm64 = run("fp64", args.fp64_steps, 1)
    fp32_leg = None
    if args.precision == "bf16x3":
        r32, ms32, _, tm32 = run("fp32", args.fp64_steps, 1)
        fp32_leg = {"ms_per_step": round(ms32, 3), "ri_k_ms": round(tm32["k_total_ms"], 3),
                    "ri_k_tflops": round(flops_k(nao, naux, nocc, nprop) / (tm32["k_total_ms"] * 1e-3) / 1e12, 3),
                    "rel_err_vs_fp64": (r32 - r64).abs().max().item() / r64.abs().max().item()}
        del r32
        log(f"[bench] fp32: {ms32:.2f} ms/step, K {tm32['k_total_ms']:.2f} ms")
    log(f"[bench] fp64: {ms64:.2f} ms/step, K {tm64['k_total_ms']:.2f} ms, J {tm64['j_total_ms']:.2f} ms")
    if args.precision == "fp64":
        rp, ms, wall, tm = run("fp64", args.steps, args.warmup)
    else:
        rp, ms, wall, tm = run(args.precision, args.steps, args.warmup)
    log(f"[bench] {args.precision}: {ms:.2f} ms/step, K {tm['k_total_ms']:.2f} ms, J {tm['j_total_ms']:.2f} ms, "
        f"hot {tm['k_hot_ms']:.2f} ms over {tm['k_hot_launches']} launches")
    diff = (rp - r64).abs().max().item()
    scale = r64.abs().max().item()

    tk = tm["k_total_ms"] * 1e-3
    value = fk / tk / 1e12
    hot_cfg, hot_name = None, HOT_KERNEL[args.precision]
    fused = args.precision == "bf16x3" and kfused_supported(nao, nocc)
    if fused:
        hot_cfg, hot_name = "fused", kfused_name(nao, nocc)
    elif args.precision == "bf16x3":
        hot_cfg, hot_ks = hot_choice(nao, naux, nocc, nprop)
        if hot_cfg in BF16X3_CONFIGS:
            hot_name = f"{BF16X3_CONFIGS[hot_cfg]}, split rule: {SPLIT_RULES.get(hot_ks, hot_ks)}"
    # dominant kernel: the hot K contraction; its algorithmic work per aux index is
    # 2 nao^2 nocc nprop (the contraction with B_P), plus 2 nao nocc^2 nprop for the fused kernel,
    # which also generates its operand Q_aP = bra_a O_P.
    hot_flops_call = 2.0 * nao**2 * nocc * nprop * naux
    if fused:
        hot_flops_call += 2.0 * nao * nocc**2 * nprop * naux
    nl = max(tm["k_hot_launches"], 1)
    t_launch = tm["k_hot_ms"] / nl * 1e-3
    achieved = hot_flops_call / nl / t_launch / 1e12
    peak = PEAK_TFLOPS[args.precision]
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 (split-bf16 x3 MFMA, fp32 accumulate, fp64 output)",
                  "fp32": "fp32", "fp64": "f64"}[args.precision],
        "data": "synthetic (seeded SplitMix64/Box-Muller cderi, QR MO coefficients, N(0,1) mo1)",
        "config": {"workload": "RI-JK response (J+K, all perturbations) on synthetic tensors",
                   "naux": naux, "nao": nao, "nocc": nocc, "nvir": nao - nocc, "nprop": nprop,
                   "precision": args.precision},
        "ri_k_tflops": round(value, 3),
        "ri_k_ms": round(tm["k_total_ms"], 3),
        "ri_j_ms": round(tm["j_total_ms"], 3),
        "fp64": {"ms_per_step": round(ms64, 3), "ri_k_ms": round(tm64["k_total_ms"], 3),
                 "ri_k_tflops": round(fk / (tm64["k_total_ms"] * 1e-3) / 1e12, 3)},
        "fp32": fp32_leg,
        "speedup_vs_fp64": round(ms64 / ms, 3),
        "max_abs_err_vs_fp64": diff,
        "rel_err_vs_fp64": diff / scale,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4),
                     "traffic": MEASURED_TRAFFIC.get((args.precision, nao, naux, nocc, nprop, hot_cfg)),
                     "counters": MEASURED_COUNTERS.get((args.precision, nao, naux, nocc, nprop, hot_cfg)),
                     "kernel": hot_name + (" (fused K term 1: operand generation + contraction, all perturbations in one launch)"
                                           if fused else " (hot K contraction, one perturbation chunk per launch)"),
                     "launches_per_step": nl, "avg_launch_ms": round(t_launch * 1e3, 4)},
    }
    if args.precision in ("bf16x3", "fp64"):
        kind = 2 if fused else (0 if args.precision == "bf16x3" else 1)
        mp = measured_mfma_peak(args.precision, kind)
        scale = 3.0 if args.precision == "bf16x3" else 1.0  # MFMA flops per algorithmic flop
        line["roofline"]["measured_mfma_peak"] = {
            "instruction": {0: "v_mfma_f32_16x16x32_bf16", 1: "v_mfma_f64_16x16x4f64",
                            2: "v_mfma_f32_32x32x16_bf16"}[kind],
            "tflops": round(mp, 1), "mfma_flops_per_algorithmic_flop": scale,
            "frac": round(achieved * scale / mp, 4)}
        # the matrix-core work the launch issues (bf16 products, padding included) against the
        # published peak: the fraction of the chip's MFMA rate the kernel sustains
        line["roofline"]["mfma_work_frac_of_peak"] = round(achieved * scale / peak, 4)
    # resident response matrix: build once per orbital set, then one HBM-bound product per call
    line["resident"] = resident_leg(d_cderi, d_C, occ, d_bra, r64, nao, nocc, nprop, args.steps)
    if not args.no_cphf:
        line["cphf"] = cphf_legs(obj, d_cderi, d_C, occ, nao, nocc, nprop, dev)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cderi, C, occ, bra, args.cpu_budget)
        if "cphf" in line:
            # the plain FP64 solve's operator applications at the oracle's measured RI-K rate
            # (RI-J and the Krylov algebra left out: a lower bound on the host's wall time)
            calls = line["cphf"]["fp64"]["aop_calls"] or 0
            per_call = flops_k(nao, naux, nocc, nprop) / (line["cpu_baseline"]["value"] * 1e12)
            line["cphf"]["cpu_oracle_estimate"] = {
                "wall_s": round(calls * per_call, 1), "aop_calls": calls, "s_per_call": round(per_call, 2),
                "cores": line["cpu_baseline"]["cores"],
                "basis": "RI-K flops per response call / cpu_baseline TFLOP/s x the FP64 leg's response calls",
                "speedup_fp64_resident": round(calls * per_call / line["cphf"]["fp64_resident"]["wall_s"], 1),
                "speedup_mixed_resident": round(calls * per_call / line["cphf"]["mixed_resident"]["wall_s"], 1)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
