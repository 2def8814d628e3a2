"""GPU: the strided batched split-K MFMA GEMM (hh_util_gemm) against a torch FP64 product."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def operands(M, N, K, ak, bk, nb0, nb1, dev, gen, pad=3):
    lda = (K if ak else M) + pad
    ldb = (K if bk else N) + pad
    ldc = M + pad
    rowsA = M if ak else K
    rowsB = N if bk else K
    nb = nb0 * nb1
    Am = torch.randn(nb, rowsA, lda, generator=gen, dtype=torch.float64).to(dev)
    Bm = torch.randn(nb, rowsB, ldb, generator=gen, dtype=torch.float64).to(dev)
    Cm = torch.randn(nb, N, ldc, generator=gen, dtype=torch.float64).to(dev)
    A = Am[:, :, :K] if ak else Am[:, :, :M].transpose(1, 2)
    B = Bm[:, :, :K].transpose(1, 2) if bk else Bm[:, :, :N]
    C = Cm[:, :, :M].transpose(1, 2)
    strides = [rowsA * lda, nb0 * rowsA * lda, rowsB * ldb, nb0 * rowsB * ldb, N * ldc, nb0 * N * ldc]
    return Am, Bm, Cm, A, B, C, lda, ldb, ldc, strides


def run(M, N, K, ak, bk, nb0=1, nb1=1, alpha=1.0, beta=0.0, prec=0, dev=None, seed=0):
    from hesship import _lib
    gen = torch.Generator().manual_seed(seed)
    Am, Bm, Cm, A, B, C, lda, ldb, ldc, strides = operands(M, N, K, ak, bk, nb0, nb1, dev, gen)
    ref = alpha * torch.matmul(A, B) + beta * C
    s6 = (ctypes.c_int64 * 6)(*strides)
    _lib.check(_lib.lib().hh_util_gemm(M, N, K, alpha, Am.data_ptr(), lda, int(ak), Bm.data_ptr(), ldb, int(bk),
                                       beta, Cm.data_ptr(), ldc, nb0, nb1, s6, prec,
                                       torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    got = Cm[:, :, :M].transpose(1, 2)
    scale = (A.abs() @ B.abs()).max().item() * abs(alpha) + 1e-300
    return (got - ref).abs().max().item() / scale


@pytest.mark.parametrize("ak", [False, True])
@pytest.mark.parametrize("bk", [False, True])
def test_gemm_layouts_fp64(dev, ak, bk):
    err = run(37, 29, 45, ak, bk, nb0=2, nb1=3, alpha=0.7, beta=-1.3, dev=dev)
    assert err < 1e-14


@pytest.mark.parametrize("shape", [(1, 1, 1), (16, 16, 4), (130, 257, 19), (300, 90, 5000), (400, 1000, 64)])
def test_gemm_shapes_fp64(dev, shape):
    M, N, K = shape
    err = run(M, N, K, False, True, dev=dev, beta=0.5)
    assert err < 1e-14


def test_gemm_zero_k(dev):
    # K = 0: C = beta C
    err = run(20, 20, 0, False, True, beta=2.0, dev=dev)
    assert err < 1e-300 or err == 0.0


@pytest.mark.parametrize("prec,tol", [(1, 2e-6), (2, 2e-5)])
@pytest.mark.parametrize("ak,bk", [(False, True), (True, False)])
def test_gemm_mixed_precision(dev, prec, tol, ak, bk):
    err = run(200, 150, 3000, ak, bk, nb0=2, beta=1.0, prec=prec, dev=dev)
    assert err < tol, err
    # and it is genuinely reduced precision (not silently FP64)
    assert err > 1e-12


# ---- NT product (hh_util_gemm_nt): tile configurations and the autotuner ----

def run_nt(M, N, K, config, ksplit, alpha=1.0, beta=0.0, dev=None, seed=0, pad=8):
    from hesship import _lib
    gen = torch.Generator().manual_seed(seed)
    lda = ldb = K + pad
    A = torch.randn(M, lda, generator=gen, dtype=torch.float64).to(dev)
    B = torch.randn(N, ldb, generator=gen, dtype=torch.float64).to(dev)
    ldc = M + 3
    Cm = torch.randn(N, ldc, generator=gen, dtype=torch.float64).to(dev)
    ref = alpha * (A[:, :K] @ B[:, :K].T) + beta * Cm[:, :M].T
    _lib.check(_lib.lib().hh_util_gemm_nt(M, N, K, alpha, A.data_ptr(), lda, B.data_ptr(), ldb, beta, Cm.data_ptr(),
                                          ldc, config, ksplit, torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    got = Cm[:, :M].T
    scale = (A[:, :K].abs() @ B[:, :K].abs().T).max().item() * abs(alpha) + 1e-300
    return (got - ref).abs().max().item() / scale


@pytest.mark.parametrize("config", [0, 1, 2])
@pytest.mark.parametrize("ksplit", [1, 3, 0, -1])
@pytest.mark.parametrize("shape", [(400, 300, 1024), (192, 1000, 2048), (37, 29, 96), (576, 576, 640)])
def test_gemm_nt_configs(dev, config, ksplit, shape):
    M, N, K = shape
    assert run_nt(M, N, K, config, ksplit, alpha=0.7, beta=-1.3, dev=dev) < 1e-13


def test_gemm_nt_autotuner(dev):
    import ctypes as C
    from hesship import _lib
    L = _lib.lib()
    cfg, ks = C.c_int32(), C.c_int32()
    M, N, K = 208, 640, 3072
    _lib.check(L.hh_util_gemm_nt_choice(M, N, K, C.byref(cfg), C.byref(ks)))
    assert cfg.value == -1
    assert run_nt(M, N, K, -1, 0, beta=0.5, dev=dev) < 1e-13      # tunes, then the product itself
    _lib.check(L.hh_util_gemm_nt_choice(M, N, K, C.byref(cfg), C.byref(ks)))
    assert cfg.value in (0, 1, 2) and ks.value in (0, -1)
    first = (cfg.value, ks.value)
    assert run_nt(M, N, K, -1, 0, beta=0.5, dev=dev, seed=1) < 1e-13  # cached choice
    _lib.check(L.hh_util_gemm_nt_choice(M, N, K, C.byref(cfg), C.byref(ks)))
    assert (cfg.value, ks.value) == first
    # a shape of the same class (N, K rounded up to powers of two) shares the choice
    _lib.check(L.hh_util_gemm_nt_choice(M, 600, 2100, C.byref(cfg), C.byref(ks)))
    assert (cfg.value, ks.value) == first
    # with the tuner off a new class keeps the default configuration and is not cached
    _lib.check(L.hh_util_set_autotune(0))
    try:
        assert run_nt(160, 96, 512, -1, 0, dev=dev) < 1e-13
        _lib.check(L.hh_util_gemm_nt_choice(160, 96, 512, C.byref(cfg), C.byref(ks)))
        assert cfg.value == -1
    finally:
        _lib.check(L.hh_util_set_autotune(1))
    with pytest.raises(_lib.HesshipError, match="configuration"):
        _lib.check(L.hh_util_gemm_nt(32, 32, 32, 1.0, None, 40, None, 40, 0.0, None, 32, 5, 0,
                                     torch.cuda.current_stream().cuda_stream))


def run_nt_bf16x3(M, N, K, config, ksplit, alpha=1.0, beta=0.0, dev=None, seed=0, pad=8):
    from hesship import _lib
    gen = torch.Generator().manual_seed(seed)
    lda = ldb = K + pad
    A = torch.randn(M, lda, generator=gen, dtype=torch.float64).to(dev)
    B = torch.randn(N, ldb, generator=gen, dtype=torch.float64).to(dev)
    ldc = M + 3
    Cm = torch.randn(N, ldc, generator=gen, dtype=torch.float64).to(dev)
    ref = alpha * (A[:, :K] @ B[:, :K].T) + beta * Cm[:, :M].T
    _lib.check(_lib.lib().hh_util_gemm_nt_bf16x3(M, N, K, alpha, A.data_ptr(), lda, B.data_ptr(), ldb, beta,
                                                 Cm.data_ptr(), ldc, config, ksplit,
                                                 torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    got = Cm[:, :M].T
    scale = (A[:, :K].abs() @ B[:, :K].abs().T).max().item() * abs(alpha) + 1e-300
    return (got - ref).abs().max().item() / scale


@pytest.mark.parametrize("config", [0, 1, 2, 3])
@pytest.mark.parametrize("ksplit", [1, 3, 0, -1])
@pytest.mark.parametrize("shape", [(400, 1800, 2048), (37, 29, 96), (208, 256, 512), (417, 300, 640),
                                   (801, 130, 320)])
def test_gemm_nt_bf16x3_configs(dev, config, ksplit, shape):
    """Split-BF16 x3 NT product, every tile configuration (208 x 256 on 1 x 8 and 2 x 4 waves, 400 x 128
    and 128 x 128 with swizzled LDS rows; M up to three 400-row tiles): error relative to
    |A| |B| at the scheme's level (missing lo*lo term ~2^-16, FP32 accumulation)."""
    M, N, K = shape
    assert run_nt_bf16x3(M, N, K, config, ksplit, alpha=0.7, beta=-1.3, dev=dev) < 3e-5


def test_gemm_nt_bf16x3_autotuner(dev):
    import ctypes as C
    from hesship import _lib
    L = _lib.lib()
    cfg, ks = C.c_int32(), C.c_int32()
    M, N, K = 400, 1000, 4096
    _lib.check(L.hh_util_gemm_nt_bf16x3_choice(M, N, K, C.byref(cfg), C.byref(ks)))
    assert cfg.value == -1
    assert run_nt_bf16x3(M, N, K, -1, 0, dev=dev) < 3e-5
    _lib.check(L.hh_util_gemm_nt_bf16x3_choice(M, N, K, C.byref(cfg), C.byref(ks)))
    assert cfg.value in (0, 1, 2, 3) and ks.value in (0, -1)
    # the F64 classes are separate
    _lib.check(L.hh_util_gemm_nt_choice(M, N, K, C.byref(cfg), C.byref(ks)))
    assert cfg.value == -1


def test_mfma_peak_probe(dev):
    """The MFMA issue-rate probe returns a positive rate per instruction, bf16 above f64."""
    import ctypes as C
    from hesship import _lib
    out = {}
    for kind, iters in ((0, 50), (1, 500), (2, 25)):
        v = C.c_double()
        _lib.check(_lib.lib().hh_util_mfma_peak(kind, iters, C.byref(v), torch.cuda.current_stream().cuda_stream))
        out[kind] = v.value
    assert out[0] > 0 and out[1] > 0 and out[2] > 0 and out[0] > out[1]
    with pytest.raises(_lib.HesshipError):
        _lib.check(_lib.lib().hh_util_mfma_peak(3, 10, C.byref(C.c_double()), None))
