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
Two builders of the same hierarchy: `build_hierarchy_native` runs every level in the library's
multithreaded host code on 3x3 / 6x6 blocks (csrc/sa_host.cpp: graph, aggregation,
fit_candidates, smoothing and the Galerkin product as block SpGEMMs) and is the one the solvers
use; `build_hierarchy` does the sparse products with scipy and is kept as the independent route
the tests compare it with. PyAMG itself is not available here; both builders equal an
independent restatement of its published algorithm (aggregates identical, P and the coarse
matrices to 1e-12), which is pinned by PyAMG's own documented examples and closed forms.
"""
import ctypes as C
import os

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
    """Node graph: one edge per stored nonzero off-diagonal (bs x bs) block of A (CSR ptr
    int64, idx int32); the strength of connection with theta = 0."""
    Ab = sp.bsr_matrix(A, blocksize=(bs, bs))
    Ab.sort_indices()
    n = Ab.shape[0] // bs
    rows = np.repeat(np.arange(n), np.diff(Ab.indptr))
    keep = (rows != Ab.indices) & (np.abs(Ab.data).reshape(len(Ab.indices), -1).max(1) > 0)
    ptr = np.zeros(n + 1, np.int64)
    ptr[1:] = np.cumsum(np.bincount(rows[keep], minlength=n))
    return ptr, Ab.indices[keep].astype(np.int32)


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


_RHO0 = [np.empty(0)]


def rho_start(n):
    """default_rng(0).standard_normal(n), the start vector of AMG._rho (gfem/linalg.py:140), as a
    read-only prefix of one cached longer draw: the generator produces its values in order, so a
    prefix of a longer draw equals the shorter draw (a 1M-dof model's draw costs ~20 ms)."""
    v = _RHO0[0]
    if len(v) < n:
        v = np.random.default_rng(0).standard_normal(max(n, 2 * len(v)))
        v.flags.writeable = False
        _RHO0[0] = v
    return v[:n]


def spectral_radius_dinvA(A, dinv, it=15):
    """1.1 x 15-step power estimate of rho(D^-1 A) from default_rng(0) (gfem/linalg.py:138-146)."""
    x = rho_start(A.shape[0]).copy()
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


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def build_hierarchy_native(A, B, bs=3, max_coarse=300, max_levels=10, omega=4.0 / 3.0, log=None,
                           nthreads=None):
    """build_hierarchy's levels from the library's host SA level builder (exfem_sa_level_build)
    on block-sparse matrices; rho from exfem_sa_rho with the same default_rng(0) start vector.
    Returns (levels, A_coarse) with BSR level matrices."""
    import time
    lib = L.load()
    nt = nthreads or _threads()
    Ab = A if (sp.isspmatrix_bsr(A) and A.blocksize == (bs, bs)) else sp.bsr_matrix(A, blocksize=(bs, bs))
    Ab.sort_indices()
    B = np.ascontiguousarray(B, dtype=np.float64)
    levels = []
    while len(levels) + 1 < max_levels and Ab.shape[0] // bs > max_coarse:
        t0 = time.time()
        n = Ab.shape[0] // bs
        K = B.shape[1]
        ptr = np.ascontiguousarray(Ab.indptr, dtype=np.int64)
        idx = np.ascontiguousarray(Ab.indices, dtype=np.int32)
        vals = np.ascontiguousarray(Ab.data, dtype=np.float64)
        h = C.c_void_p()
        L.check(lib.exfem_sa_level_build(n, bs, ptr.ctypes.data, idx.ctypes.data, vals.ctypes.data,
                                         K, B.ctypes.data, omega, 1e-10, nt, C.byref(h)))
        try:
            nagg, pn, cn = C.c_int64(), C.c_int64(), C.c_int64()
            L.check(lib.exfem_sa_level_info(h, C.byref(nagg), C.byref(pn), C.byref(cn)))
            nagg, pn, cn = nagg.value, pn.value, cn.value
            if nagg == 0 or nagg >= n:
                break
            agg = np.empty(n, np.int32)
            p_ptr, p_idx, p_val = np.empty(n + 1, np.int64), np.empty(max(pn, 1), np.int32), np.empty((max(pn, 1), bs, K))
            c_ptr, c_idx, c_val = np.empty(nagg + 1, np.int64), np.empty(max(cn, 1), np.int32), np.empty((max(cn, 1), K, K))
            Bc = np.empty((nagg * K, K))
            L.check(lib.exfem_sa_level_get(h, agg.ctypes.data, p_ptr.ctypes.data, p_idx.ctypes.data,
                                           p_val.ctypes.data, c_ptr.ctypes.data, c_idx.ctypes.data,
                                           c_val.ctypes.data, Bc.ctypes.data))
        finally:
            lib.exfem_sa_level_free(h)
        P = sp.bsr_matrix((p_val[:pn], p_idx[:pn], p_ptr), shape=(n * bs, nagg * K))
        Ac = sp.bsr_matrix((c_val[:cn], c_idx[:cn], c_ptr), shape=(nagg * K, nagg * K))
        d = Ab.diagonal()
        dinv = np.where(np.abs(d) > 0, 1.0 / np.where(d == 0, 1, d), 0.0)
        x = rho_start(n * bs).copy()
        rho = C.c_double()
        L.check(lib.exfem_sa_rho(n, bs, ptr.ctypes.data, idx.ctypes.data, vals.ctypes.data,
                                 x.ctypes.data, 15, nt, C.byref(rho)))
        levels.append(dict(A=Ab, P=P, R=P.T.tobsr(), dinv=dinv, rho=rho.value, n=n * bs,
                           nc=nagg * K, nagg=nagg, agg=agg, B=B))
        if log:
            log(f"SA level {len(levels) - 1}: {n * bs} rows, nnz {Ab.nnz} -> {nagg * K} rows "
                f"({time.time() - t0:.1f}s)")
        Ab, B, bs = Ac, Bc, K
    return levels, Ab


def bc_bsr(A, free, bs=3):
    """bc_matrix (gfem/linalg.py:213-217) on a BSR matrix: D A D + (I - D) with D = diag(free),
    without leaving the block format."""
    A = A.copy()
    n = A.shape[0] // bs
    f = np.asarray(free, dtype=np.float64).reshape(n, bs)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    if np.count_nonzero(rows == A.indices) != n:          # a node without a diagonal block
        Dm = sp.diags(f.ravel())
        M = (Dm @ A.tocsr() @ Dm + sp.diags(1.0 - f.ravel())).tocsr()
        return sp.bsr_matrix(M, blocksize=(bs, bs))
    A.data *= f[rows][:, :, None] * f[A.indices][:, None, :]
    diag = rows == A.indices
    A.data[diag] += np.einsum("kr,rc->krc", 1.0 - f[rows[diag]], np.eye(bs))
    return A


def csr_arrays(M):
    M = sp.csr_matrix(M)
    M.sort_indices()
    if M.nnz >= 2 ** 31:
        raise ValueError("level matrix too large for int32 indices")
    return (np.ascontiguousarray(M.indptr.astype(np.int32)),
            np.ascontiguousarray(M.indices.astype(np.int32)),
            np.ascontiguousarray(M.data.astype(np.float64)))
