"""Helpers for checking full-size device results against the oracle on a
sample of elements.

The reference right-hand side is element-local
(src/physics/primitive_equations.cpp:428-442: each element's tendencies depend
on that element's state and bathymetry only), and so is every Runge-Kutta
stage built from it.  The sampled elements' rows of a state are therefore a
self-contained input: the oracle run on them alone gives exactly the values
the full mesh gives for those elements.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def sample_elements(E, k=400, seed=0):
    """k distinct elements chosen at random, plus the first and the last."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(E, min(k, E), replace=False)
    return np.unique(np.r_[0, idx, E - 1]).astype(np.int64)


def rows(a, idx, width):
    """Rows idx of an element-major host array with `width` values per element."""
    return np.ascontiguousarray(np.asarray(a).reshape(-1, width)[idx].reshape(-1))


def lines_rows(T, idx, E, n):
    """Packed rows ([e][Hu|Hv|eta|HT|HS][ndof], float64, host) of the elements
    idx of a device state in the line layout [field][i][e][j + n k]."""
    import torch

    it = torch.from_numpy(idx).to(T.device)
    V = T.view(5, n, E, n * n).index_select(2, it)  # (5, n, S, n^2)
    return np.ascontiguousarray(
        V.permute(2, 0, 3, 1).reshape(-1).double().cpu().numpy())


def oracle_chunks(fn, U, h, nd, threads=16):
    """fn(U_chunk, h_chunk) -> packed output, run over element chunks on a
    thread pool (the oracle's C code holds no global state and ctypes releases
    the interpreter lock), then concatenated in element order."""
    E = U.size // (5 * nd)
    bounds = np.linspace(0, E, min(threads, E) + 1).astype(int)
    parts = [(bounds[i], bounds[i + 1]) for i in range(len(bounds) - 1)
             if bounds[i + 1] > bounds[i]]

    def one(p):
        e0, e1 = p
        return fn(np.ascontiguousarray(U[e0 * 5 * nd:e1 * 5 * nd]),
                  np.ascontiguousarray(h[e0 * nd:e1 * nd]))

    with ThreadPoolExecutor(len(parts)) as ex:
        outs = list(ex.map(one, parts))
    return np.concatenate(outs)


def rel_l2_fields(a, b, nd):
    """Relative L2 error of a against b per field of a packed state."""
    A = np.asarray(a).reshape(-1, 5, nd)
    B = np.asarray(b).reshape(-1, 5, nd)
    out = []
    for f in range(5):
        den = np.linalg.norm(B[:, f])
        out.append(float(np.linalg.norm(A[:, f] - B[:, f]) / (den if den > 0 else 1.0)))
    return out
