"""The allocation-free dense solve (hh_linalg_solve_noalloc) against the allocating one
(hh_linalg_solve, the projected solve of krylov_block.rs:214) through the C ABI: two calls in a
row, the second smaller, each compared with the allocating entry point and checked by its residual;
sizes beyond the static scratch are rejected."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _system(nd, nrhs, seed, dev):
    rng = np.random.default_rng(seed)
    H = rng.standard_normal((nd, nd)) + nd ** 0.5 * np.eye(nd)
    G = rng.standard_normal((nd, nrhs))
    f = lambda x: torch.from_numpy(np.asfortranarray(x).T.copy()).to(dev).T  # column-major
    return H, G, f


def _solve(entry, H, G, f):
    from hesship import _lib
    Hd, Gd = f(H), f(G)
    nd, nrhs = G.shape
    _lib.check(getattr(_lib.lib(), entry)(Hd.data_ptr(), nd, Gd.data_ptr(), nrhs,
                                          torch.cuda.current_stream().cuda_stream))
    return Gd.cpu().numpy()


@pytest.mark.parametrize("order", ["noalloc_first", "regular_first"])
def test_noalloc_solve_matches_allocating_solve(dev, order):
    for nd, nrhs, seed in ((300, 7, 1), (70, 3, 2)):   # the second call smaller than the first
        H, G, f = _system(nd, nrhs, seed, dev)
        if order == "noalloc_first":
            x0 = _solve("hh_linalg_solve_noalloc", H, G, f)
            x1 = _solve("hh_linalg_solve", H, G, f)
        else:
            x1 = _solve("hh_linalg_solve", H, G, f)
            x0 = _solve("hh_linalg_solve_noalloc", H, G, f)
        err = np.abs(x0 - x1).max() / np.abs(x1).max()
        res = np.abs(H @ x0 - G).max() / np.abs(G).max()
        print(f"noalloc solve nd={nd} nrhs={nrhs}: rel diff {err:.2e}, residual {res:.2e}")
        assert err < 1e-10 and res < 1e-10


def test_noalloc_solve_rejects_sizes_beyond_scratch(dev):
    from hesship import _lib
    H = torch.zeros(4, dtype=torch.float64, device=dev)
    with pytest.raises(_lib.HesshipError, match="scratch"):
        _lib.check(_lib.lib().hh_linalg_solve_noalloc(H.data_ptr(), 8193, H.data_ptr(), 1,
                                                      torch.cuda.current_stream().cuda_stream))
