"""GPU parity at full production sizes.

- The synthetic RI-JK workload (naux 2400, nao 400, nocc 40, nprop 90; seeded as hesship/synth.py)
  through the production launch of every precision: all 90 perturbations in one call, whatever
  tile, split and batch the library picks, and three output columns (each depends only on its
  own bra) compared with the oracle's K loop and J (hess_r.rs:82-186, 1916-1953).
- The RKS grid response at (H2O)_8 sizes in nao / nocc / nprop (192, 40, 72; GGA; 128-point grid
  batches) on 8192 grid points, three columns against the oracle (hess_rks.rs:570-605).
Tolerances relative to max |oracle|: FP64 1e-10, FP32 2e-5, BF16x3 1e-4.
"""
import time

import numpy as np
import pytest
import torch

from oracle import numint as oni
from oracle import rijk as orc

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 2e-5, "bf16x3": 1e-4}
COLS = [0, 45, 89]


@pytest.fixture(scope="module")
def c2(dev):
    from hesship import synth
    t0 = time.perf_counter()
    cderi, C, occ, bra = synth.rijk_workload(400, 2400, 40, 90)
    ref = orc.rhf_response_bra(cderi, C, occ, bra[:, :, COLS], 1.0, 1.0)
    print(f"c2 inputs and oracle columns in {time.perf_counter() - t0:.1f} s", flush=True)
    return cderi, C, occ, bra, ref


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_rijk_response_full_size_production_launch(dev, c2, precision):
    import hesship
    cderi, C, occ, bra, ref = c2
    d_cderi = torch.from_numpy(np.asfortranarray(cderi).T.copy()).to(dev).T
    obj = hesship.RHessRIJK(d_cderi, 1.0, 1.0)
    obj.make_response_preparation(torch.from_numpy(C).to(dev), occ)
    d_bra = hesship.fortran(torch.from_numpy(bra).to(dev))
    got = obj.get_response_bra(d_bra, precision=precision)
    got = got[:, :, COLS].cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"RI-JK response naux 2400 nao 400 nocc 40 nprop 90 ({precision}): rel err of columns {COLS} {err:.2e}")
    assert err < TOL[precision]


def test_rijk_response_full_size_materialized_k(dev, c2):
    """The BF16x3 K path with the operand T materialised (the path taken past the fused kernel's
    limits): the autotuned NT hot product at K = naux x 416 = 998400 over all 90 perturbations,
    three columns against the oracle."""
    import ctypes as C
    import hesship
    from hesship import _lib
    cderi, Cm, occ, bra, ref = c2
    d_cderi = torch.from_numpy(np.asfortranarray(cderi).T.copy()).to(dev).T
    obj = hesship.RHessRIJK(d_cderi, 1.0, 1.0)
    obj.set_k_path("materialized")
    obj.make_response_preparation(torch.from_numpy(Cm).to(dev), occ)
    d_bra = hesship.fortran(torch.from_numpy(bra).to(dev))
    got = obj.get_response_bra(d_bra, precision="bf16x3")[:, :, COLS].cpu().numpy()
    cfg, ks = C.c_int32(), C.c_int32()
    _lib.check(_lib.lib().hh_util_gemm_nt_bf16x3_choice(400, 3600, 2400 * 416, C.byref(cfg), C.byref(ks)))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"materialised K, hot product configuration {cfg.value} split rule {ks.value}: rel err {err:.2e}")
    assert err < TOL["bf16x3"]


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_rks_response_c3_shape(dev, precision):
    import hesship
    ng, nao, nocc, nprop = 8192, 192, 40, 72
    rng = np.random.default_rng(0x5EED0005)
    ao = rng.standard_normal((ng, nao, 4)) * 0.3
    w = rng.uniform(0.0, 0.1, ng)
    f = rng.standard_normal((ng, 4, 4)) * 0.1
    f = f + f.transpose(0, 2, 1)
    mocc = rng.standard_normal((nao, nocc)) * 0.1
    mo1 = rng.standard_normal((nao, nocc, nprop)) * 0.1
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    ni = hesship.NIMatmul(T(ao), T(w), nbatch=128)
    got = ni.get_rks_response_bra("sigma", T(f), T(mo1), T(mocc), precision=precision)[:, :, COLS[:2] + [71]]
    ref = oni.rks_response_bra("sigma", f, ao, w, mo1[:, :, COLS[:2] + [71]], mocc)
    err = np.abs(got.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"RKS response ng {ng} nao {nao} nocc {nocc} nprop {nprop} ({precision}): rel err {err:.2e}")
    assert err < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_rijk_resident_response_full_size(dev, c2, precision):
    """The resident response matrix (16000 x 16000 at these sizes) built and applied at full size."""
    import hesship
    cderi, C, occ, bra, ref = c2
    d_cderi = torch.from_numpy(np.asfortranarray(cderi).T.copy()).to(dev).T
    obj = hesship.RHessRIJK(d_cderi, 1.0, 1.0)
    obj.set_response_path("resident")
    obj.make_response_preparation(torch.from_numpy(C).to(dev), occ)
    d_bra = hesship.fortran(torch.from_numpy(bra).to(dev))
    got = obj.get_response_bra(d_bra, precision=precision)[:, :, COLS].cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"resident RI-JK response, full size ({precision}): rel err of columns {COLS} {err:.2e}")
    assert err < TOL[precision]
