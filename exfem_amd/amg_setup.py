"""Smoothed-aggregation hierarchy, built on the host (the setup the reference delegates to
PyAMG, gfem/linalg.py:103-136).

Algorithm (PyAMG's published smoothed_aggregation_solver with the reference's options
strength=("symmetric", theta=0), smooth=("jacobi", omega=4/3, weighting="local"),
max_coarse=300, improve_candidates=None, near-null space = the 6 rigid-body modes on 3x3 nodal
blocks):
  per level, while the number of block rows exceeds max_coarse:
    1. strength graph = block sparsity pattern of A (theta = 0 keeps every stored block);
    2. standard (greedy 3-pass) aggregation of that graph;
    3. tentative prolongator T and coarse candidates Bc by a per-aggregate modified
       Gram-Schmidt of the candidate rows (fit_candidates, tol 1e-10);
    4. P = T - omega D^-1 A T with D the row sums of |A| (local Gershgorin weighting);
    5. R = P^T, A_c = R A P; the coarse block size becomes the number of candidates.
Aggregation and fit_candidates run in the HIP library's host code (csrc/sa_host.cpp); the
sparse products use scipy. PyAMG itself is not available here, so the hierarchy is not
bit-identical to the reference's; it is checked through SA invariants in the tests.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L


def rigid_body_modes(X, rot_mask=None):
    """Near-null space (3 nn, 6): translations and rotations about the centroid
    (gfem/linalg.py:80-98, no rotation rows)."""
    X = np.asarray(X, float)
    n = len(X)
    c = X - X.mean(0)
    B = np.zeros((3 * n, 6))
    for i in range(3):
        B[i::3, i] = 1.0
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    B[0::3, 3], B[1::3, 3] = -y, x
    B[1::3, 4], B[2::3, 4] = -z, y
    B[0::3, 5], B[2::3, 5] = z, -x
    return B


def block_graph(A, bs):
    """Node graph: one edge per stored (bs x bs) block of A (CSR ptr int64, idx int32)."""
    Ab = sp.bsr_matrix(A, blocksize=(bs, bs))
    Ab.sort_indices()
    return Ab.indptr.astype(np.int64), Ab.indices.astype(np.int32)


def aggregate(ptr, idx, n):
    agg = np.empty(n, np.int32)
    nagg = L.load().exfem_sa_aggregate(n, ptr.ctypes.data, idx.ctypes.data, agg.ctypes.data)
    return agg, int(nagg)


def fit_candidates(agg, nagg, B, K1, tol=1e-10):
    n = len(agg)
    K2 = B.shape[1]
    B = np.ascontiguousarray(B, dtype=np.float64)
    Q = np.empty((n * K1, K2))
    R = np.empty((nagg, K2, K2))
    L.check(L.load().exfem_sa_fit_candidates(n, K1, K2, agg.ctypes.data, nagg, B.ctypes.data,
                                              tol, Q.ctypes.data, R.ctypes.data))
    rows = np.repeat(np.arange(n * K1, dtype=np.int64), K2)      # row r of Q for each (r, c)
    a = agg[rows // K1].astype(np.int64)
    cols = a * K2 + np.tile(np.arange(K2, dtype=np.int64), n * K1)
    keep = a >= 0
    T = sp.csr_matrix((Q.ravel()[keep], (rows[keep], cols[keep])), shape=(n * K1, nagg * K2))
    return T, R.reshape(nagg * K2, K2)


def jacobi_smooth(A, T, omega=4.0 / 3.0):
    """P = T - omega D^-1 A T, D = row sums of |A| (weighting="local")."""
    D = np.asarray(abs(A).sum(axis=1)).ravel()
    Dinv = np.zeros_like(D)
    nz = D != 0
    Dinv[nz] = 1.0 / D[nz]
    AT = (A @ T).tocsr()
    P = (T - sp.diags(omega * Dinv) @ AT).tocsr()
    P.sort_indices()
    return P


def spectral_radius_dinvA(A, dinv, it=15):
    """1.1 x 15-step power estimate of rho(D^-1 A) from default_rng(0) (gfem/linalg.py:138-146)."""
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    lam = 1.0
    for _ in range(it):
        y = dinv * (A @ x)
        lam = np.linalg.norm(y) / max(np.linalg.norm(x), 1e-300)
        x = y / max(np.linalg.norm(y), 1e-300)
    return 1.1 * lam


def coarse_cholesky(Ac):
    """gfem/linalg.py:126-133."""
    Ac = np.asarray(Ac, float)
    Ac = 0.5 * (Ac + Ac.T)
    try:
        return np.linalg.cholesky(Ac + 1e-13 * np.abs(Ac).max() * np.eye(len(Ac)))
    except np.linalg.LinAlgError:
        w, V = np.linalg.eigh(Ac)
        return np.linalg.cholesky(V @ np.diag(np.maximum(w, 1e-12 * w.max())) @ V.T)


def build_hierarchy(A, B, bs=3, max_coarse=300, max_levels=10, omega=4.0 / 3.0, log=None):
    """Returns (levels [dict(A, P, R, dinv, rho, n, nc, nagg)], A_coarse csr). `log(msg)`, if
    given, receives one line per finished level."""
    import time
    A = sp.csr_matrix(A)
    A.sort_indices()
    levels = []
    while len(levels) + 1 < max_levels and A.shape[0] // bs > max_coarse:
        t0 = time.time()
        nb = A.shape[0] // bs
        ptr, idx = block_graph(A, bs)
        agg, nagg = aggregate(ptr, idx, nb)
        if nagg == 0 or nagg >= nb:
            break
        T, Bc = fit_candidates(agg, nagg, B, bs)
        P = jacobi_smooth(A, T, omega)
        R = P.T.tocsr()
        R.sort_indices()
        Ac = (R @ (A @ P)).tocsr()
        Ac.sort_indices()
        d = A.diagonal()
        dinv = np.where(np.abs(d) > 0, 1.0 / np.where(d == 0, 1, d), 0.0)
        levels.append(dict(A=A, P=P, R=R, dinv=dinv, rho=spectral_radius_dinvA(A, dinv),
                           n=A.shape[0], nc=P.shape[1], nagg=nagg, agg=agg, T=T, B=B))
        if log:
            log(f"SA level {len(levels) - 1}: {A.shape[0]} rows, nnz {A.nnz} -> {Ac.shape[0]} rows "
                f"({time.time() - t0:.1f}s)")
        A, B, bs = Ac, Bc, B.shape[1]
    return levels, A


def csr_arrays(M):
    M = sp.csr_matrix(M)
    M.sort_indices()
    if M.nnz >= 2 ** 31:
        raise ValueError("level matrix too large for int32 indices")
    return (np.ascontiguousarray(M.indptr.astype(np.int32)),
            np.ascontiguousarray(M.indices.astype(np.int32)),
            np.ascontiguousarray(M.data.astype(np.float64)))
