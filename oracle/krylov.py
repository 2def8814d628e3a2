"""Oracle: block Krylov solver for (I + A) x = b.  Test infrastructure only.

Restates src/util/krylov_block.rs (ajz34/rstsr-showcase-hessian):
  krylov_block     39-197
  projected_solve  204-215
  orth_block       225-253
"""
import numpy as np


class KrylovNotConverged(RuntimeError):
    pass


def orth_block(vec, lindep):
    """krylov_block.rs:225-253: serial MGS over columns, non-normalised, drop ||v||^2 <= lindep."""
    n, nblock = vec.shape
    result, norms = [], []
    for i in range(nblock):
        vi = vec[:, i].copy()
        for j in range(len(result)):
            vi -= (vi @ result[j]) / norms[j] * result[j]
        nsq = vi @ vi
        if nsq > lindep:
            result.append(vi)
            norms.append(nsq)
    if not result:
        return np.zeros((n, 0)), []
    return np.stack(result, axis=1), norms


def projected_solve(xs, ax, inner, b):
    """krylov_block.rs:204-215: h = xs^T ax + diag(inner); c = h^-1 xs^T b; return xs c."""
    h = xs.T @ ax
    h[np.diag_indices_from(h)] += np.asarray(inner)
    g = xs.T @ b
    return xs @ np.linalg.solve(h, g)


def krylov_block(aop, b, x0=None, tol=1e-8, max_cycle=42, max_space=14, lindep=1e-14, stats=None):
    """krylov_block.rs:39-197.  Raises KrylovNotConverged where the reference panics."""
    b = np.asarray(b, dtype=np.float64)
    n, nset = b.shape
    x_accum = np.zeros((n, nset)) if x0 is None else np.array(x0, dtype=np.float64)
    conv = max(lindep, tol * tol)
    total, restart = 0, 0
    last_nd, last_b = 0, b.copy()
    last_max = 0.0
    xs = ax = None
    inner_all = []
    while total < max_cycle:
        bres = b.copy() if (restart == 0 and x0 is None) else b - (x_accum + aop(x_accum))
        x1, ip = orth_block(bres, lindep)
        if x1.shape[1] == 0:
            last_nd, last_b = 0, bres
            break
        xs_l, ax_l, inner_all = [], [], []
        nd = 0
        converged = False
        for _ in range(max_space):
            if total >= max_cycle:
                break
            total += 1
            axt = aop(x1)
            xs_l.append(x1)
            ax_l.append(axt)
            inner_all.extend(ip)
            nd += x1.shape[1]
            xs = np.concatenate(xs_l, axis=1)
            ax = np.concatenate(ax_l, axis=1)
            coeffs = (xs.T @ axt) / np.asarray(inner_all)[:, None]
            x1_new = axt - xs @ coeffs
            nx1, nip = orth_block(x1_new, lindep)
            max_ip = max(nip) if nip else 0.0
            keep = [v > conv for v in nip]
            if not all(keep):
                nx1 = nx1[:, np.asarray(keep, dtype=bool)]
                nip = [v for v, k in zip(nip, keep) if k]
            x1, ip = nx1, nip
            last_max = max_ip
            if max_ip < conv:
                converged = True
                break
        last_nd, last_b = nd, bres
        if converged or total >= max_cycle:
            break
        x_accum = x_accum + projected_solve(xs, ax, inner_all, last_b)
        restart += 1
    if last_nd == 0:
        x_out = x_accum
    else:
        x_out = x_accum + projected_solve(xs[:, :last_nd], ax[:, :last_nd], inner_all, last_b)
    if stats is not None:
        stats.update(cycles=total, restarts=restart, max_innerprod=last_max)
    if last_max >= conv:
        raise KrylovNotConverged(
            f"krylov_block failed to converge: max(||v||^2) = {last_max:.3e} >= tol^2 = {conv:.3e}")
    return x_out


def closed_form_test_system(n=40, nrhs=12):
    """The reference's deterministic Krylov fixture (prototype/02-7-krylov_testing_setup.py:5-63),
    restated: A = V diag(eigs) V^-1 with a clustered spectrum plus -0.9 and 5.0, V Hilbert-like
    + 0.5 I, b[k, j] = cos(0.7 k j) + 0.3 sin(1.1 k + 0.2 j) (k, j 1-based).  Returns (A, b[n, nrhs])."""
    eigs = np.empty(n)
    eigs[0] = -0.9
    eigs[-1] = 5.0
    eigs[1:-1] = 1.0 + 0.05 * np.cos(np.arange(n - 2) * 1.7 + 0.3)
    eigs = eigs[np.argsort(np.sin(np.arange(n) * 1.3 + 0.5))]
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    V = 1.0 / (i + j + 1.0) + 0.5 * np.eye(n)
    A = V @ np.diag(eigs) @ np.linalg.inv(V)
    k = np.arange(nrhs)[:, None] + 1.0
    c = np.arange(n)[None, :] + 1.0
    b = np.cos(0.7 * k * c) + 0.3 * np.sin(1.1 * k + 0.2 * c)
    return A, b.T.copy()
