"""Linear algebra on the MI355X path (gfem/linalg.py:21-217).

* ``pcg``: PCG with the reference recurrence and stopping rule. With an ``BcOperator`` and a
  device preconditioner it runs entirely in the HIP library (scalars resident on the device);
  with arbitrary callables it runs the same recurrence over device tensors.
* ``AMG``: smoothed-aggregation hierarchy built on the host (amg_setup.py), V-cycle on the
  device (Chebyshev smoothing, ELL level products, dense coarse solve).
* ``csr_to_device`` / ``csr_mv``, ``rigid_body_modes``, ``bc_matrix``, ``Direct``.
"""
import ctypes as C
import time

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .amg_setup import (bc_bsr, build_hierarchy, build_hierarchy_native, coarse_cholesky, csr_arrays,
                        rho_start, rigid_body_modes)
from .model import as_dev, device, stream_ptr, torch, vptr


class BcOperator:
    """A v = free*K(free*v) + (1-free)*v over a device EbeOperator (gfem/solvers.py:257-267)."""

    def __init__(self, op, matrix_free=False, dtype=L.F64):
        self.op, self.matrix_free, self.dtype = op, matrix_free, dtype

    @property
    def n(self):
        return self.op.ndof

    def __call__(self, v):
        return self.op.matvec(v, bc=True, matrix_free=self.matrix_free, dtype=self.dtype)


class Jacobi:
    """M r = pinv * r with pinv = where(free, 1/diag (1 if 0), 1) (gfem/solvers.py:432-434)."""
    kind = L.PC_JACOBI

    def __init__(self, pinv):
        self.pinv = as_dev(pinv)

    def __call__(self, r):
        return precond_apply(L.PC_JACOBI, r, pinv=self.pinv)


class Chebyshev:
    """6-step Chebyshev-Jacobi preconditioner over [lmax/30, lmax] (gfem/solvers.py:289-307) on
    the masked operator A (a BcOperator); each application runs in the library
    (exfem_pcg_precond). fp32=True runs the polynomial in FP32 inside an FP64 solve
    (EXFEM_PC_CHEBY_MIXED): FP32 element matrices attached to A's operator, or a matrix-free A."""

    def __init__(self, dinv, lmax, A, fp32=False):
        self.pinv, self.lmax, self.A = as_dev(dinv), float(lmax), A
        if fp32 and A.dtype != L.F64:
            raise ValueError("the FP32 Chebyshev polynomial preconditions FP64 solves")
        self.kind = L.PC_CHEBY_MIXED if fp32 else L.PC_CHEBY

    def __call__(self, r):
        return precond_apply(self.kind, r, op=self.A.op, pinv=self.pinv, lmax=self.lmax,
                             matrix_free=self.A.matrix_free, dtype=self.A.dtype)


def precond_apply(pc, r, op=None, pinv=None, lmax=0.0, amg=None, matrix_free=False, dtype=None):
    """z = M r, one application of the preconditioner `pc` (EXFEM_PC_*) through the library
    (exfem_pcg_precond): the M of _pcg_fn, gfem/solvers.py:286-316."""
    if dtype is None:
        dtype = L.F32 if isinstance(r, torch().Tensor) and r.dtype == torch().float32 else L.F64
    rd = as_dev(r, dtype)
    z = torch().empty_like(rd)
    ws = _workspace(rd.numel())
    L.check(L.load().exfem_pcg_precond(ws.h, None if op is None else op.h, pc, vptr(pinv),
                                       float(lmax), None if amg is None else amg.h,
                                       int(matrix_free), vptr(rd), vptr(z), dtype, stream_ptr()))
    return z


def lmax_power(A, dinv, iters=25):
    """_LinearSystem._lmax (gfem/solvers.py:328-342) on the device: 1.1 x the `iters`-th power
    ratio |D^-1 A v| / |v| of the masked operator A (a BcOperator, FP64) from
    v = sin(0.37 i + 0.1) free + 1e-3 (exfem_pcg_lmax)."""
    ws = _workspace(A.n)
    lam = C.c_double()
    L.check(L.load().exfem_pcg_lmax(ws.h, A.op.h, vptr(as_dev(dinv)), int(A.matrix_free),
                                    int(iters), C.byref(lam), stream_ptr()))
    return lam.value


class AMGPrecond:
    """M r = free * V(free * r) + (1 - free) * r with level 0 on the EbE operator
    (gfem/solvers.py:308-314). dtype: precision of the vectors (the solve); cycle: precision of
    the V-cycle (default: dtype). cycle F32 with dtype F64 is the FP32-cycle preconditioner of an
    FP64 solve (L.PC_AMG_MIXED): level 0 then needs FP32 element matrices attached to the operator
    (or the matrix-free form)."""

    def __init__(self, amg, A=None, dinv0=None, dtype=L.F64, cycle=None):
        self.amg, self.A, self.dtype = amg, A, dtype
        self.cycle = dtype if cycle is None else cycle
        if self.cycle == L.F64 and dtype == L.F32:
            raise ValueError("an FP64 cycle on FP32 vectors is not a supported form")
        self.kind = L.PC_AMG_MIXED if self.cycle != dtype else L.PC_AMG
        self.pinv = None if dinv0 is None else as_dev(dinv0)
        if self.cycle == L.F32:
            amg.set_precision(L.F32)

    def __call__(self, r):
        if self.kind == L.PC_AMG_MIXED:
            return self.amg.apply(as_dev(r, L.F64).float(), A0=self.A, dinv0=self.pinv,
                                  dtype=L.F32).double()
        return self.amg.apply(r, A0=self.A, dinv0=self.pinv, dtype=self.dtype)

    def multi(self, R):
        """The preconditioner on every row of R (m, n): batched V-cycles of up to 64 rows."""
        return self.amg.apply_multi(R, A0=self.A, dinv0=self.pinv, dtype=self.dtype,
                                    cycle=self.cycle)


_WS = {}
_USE_GRAPH = [True]


def pcg_graphs(enable=True):
    """Replay PCG iterations as captured HIP graphs (default) or launch kernels one by one."""
    _USE_GRAPH[0] = bool(enable)


def _workspace(n):
    ws = _WS.get(n)
    if ws is None:
        h = C.c_void_p()
        L.check(L.load().exfem_pcg_create(C.byref(h), n))
        ws = _PcgWorkspace(h)
        _WS.clear()
        _WS[n] = ws
    if ws.graph != _USE_GRAPH[0]:
        L.check(L.load().exfem_pcg_set_graph(ws.h, int(_USE_GRAPH[0])))
        ws.graph = _USE_GRAPH[0]
    return ws


class _PcgWorkspace:
    def __init__(self, h):
        self.h = h
        self.graph = True

    def __del__(self):
        if self.h is not None and L is not None and L._lib is not None:
            try:
                L._lib.exfem_pcg_destroy(self.h)
            except Exception:
                pass
            self.h = None


def pcg_device(A, b, M=None, tol=1e-8, maxiter=2000):
    """Fused device PCG: A a BcOperator, M None / Jacobi / Chebyshev / AMGPrecond.
    Returns (x, k, relres, status)."""
    dt = A.dtype
    bd = as_dev(b, dt)
    x = torch().empty_like(bd)
    pc = L.PC_NONE if M is None else M.kind
    pinv = None if M is None else M.pinv
    lmax = getattr(M, "lmax", 0.0)
    amg = M.amg.h if pc in (L.PC_AMG, L.PC_AMG_MIXED) else None
    if amg is not None and M.A is None:
        raise ValueError("the fused AMG-PCG needs level 0 on the EbE operator")
    ws = _workspace(A.n)
    it = C.c_int()
    rr = C.c_double()
    st = L.check(L.load().exfem_pcg_solve(ws.h, A.op.h, vptr(bd), vptr(x), dt, pc, vptr(pinv),
                                          lmax, amg, int(A.matrix_free), tol, int(maxiter),
                                          C.byref(it), C.byref(rr), stream_ptr()))
    return x, it.value, rr.value, st


def pcg(A, b, M=None, x0=None, tol=1e-8, maxiter=2000):
    """gfem/linalg.py:21-47 -> (x, k, relres)."""
    if isinstance(A, BcOperator) and x0 is None and (M is None or hasattr(M, "kind")):
        x, k, rr, _ = pcg_device(A, b, M, tol, maxiter)
        return x, k, rr
    # arbitrary callables: the same recurrence over device tensors
    M = M or (lambda r: r)
    b = as_dev(b) if not isinstance(b, torch().Tensor) else b
    x = torch().zeros_like(b) if x0 is None else as_dev(x0, L.F64 if b.dtype == torch().float64 else L.F32)
    r = b - A(x)
    z = M(r)
    p = z
    rz = torch().dot(r, z)
    bn2 = torch().clamp(torch().dot(b, b), min=1e-300)
    tol2 = tol * tol * bn2
    rr = torch().dot(r, r)
    k = 0
    while bool(rr > tol2) and k < maxiter:
        Ap = A(p)
        alpha = rz / torch().dot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        rr = torch().dot(r, r)
        z = M(r)
        rz_new = torch().dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
        k += 1
    return x, k, float(torch().sqrt(rr / bn2))


# ---------------------------------------------------------------------------------- CSR
def csr_to_device(A, ell_max_ratio=2.0, dtype=L.F64):
    """Device sparse matrix (gfem/linalg.py:51-70): ELL when the longest row is at most
    ell_max_ratio x the mean (entries k-major: entry k of row i at k*n + i, so a wavefront reads
    consecutive rows), else CSR. int32 indices; values in dtype."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    cnt = np.diff(A.indptr)
    n = A.shape[0]
    w = int(cnt.max()) if n else 0
    dev = device()
    if n and w <= ell_max_ratio * max(cnt.mean(), 1.0):
        pos = np.arange(A.nnz) - np.repeat(A.indptr[:-1], cnt)
        rows = np.repeat(np.arange(n), cnt)
        cols = np.zeros((w, n), np.int32)
        data = np.zeros((w, n), np.float64)
        cols[pos, rows] = A.indices
        data[pos, rows] = A.data
        return dict(ell_data=as_dev(data.ravel(), dtype),
                    ell_cols=torch().as_tensor(cols.ravel(), device=dev), ell_w=w, n=n, nnz=A.nnz,
                    dtype=dtype)
    return dict(data=as_dev(A.data, dtype),
                cols=torch().as_tensor(A.indices.astype(np.int32), device=dev),
                ptr=torch().as_tensor(A.indptr.astype(np.int32), device=dev), n=n, nnz=A.nnz,
                dtype=dtype)


def csr_mv(M, x, n):
    """y = M x (gfem/linalg.py:73-77) through the library (exfem_csr_mv): every row sums its
    entries in stored order, no atomics."""
    dt = M["dtype"]
    xd = as_dev(x, dt)
    y = torch().empty(M["n"], dtype=xd.dtype, device=xd.device)
    if "ell_data" in M:
        args = (M["ell_w"], None, vptr(M["ell_cols"]), vptr(M["ell_data"]))
    else:
        args = (-1, vptr(M["ptr"]), vptr(M["cols"]), vptr(M["data"]))
    L.check(L.load().exfem_csr_mv(M["n"], M["nnz"], *args, vptr(xd), vptr(y), dt, stream_ptr()))
    return y


def bc_matrix(A, free):
    """D A D + (I - D) (gfem/linalg.py:213-217)."""
    D = sp.diags(np.asarray(free, float))
    return (D @ A @ D + sp.diags(1.0 - np.asarray(free, float))).tocsr()


class Direct:
    """Host sparse LU (gfem/linalg.py:199-210) - the CPU direct solve of the reference."""

    def __init__(self, A):
        import scipy.sparse.linalg as spla
        self.n = A.shape[0]
        try:
            self._solve = spla.splu(sp.csc_matrix(A), permc_spec="COLAMD").solve
        except RuntimeError:
            self._solve = spla.factorized(sp.csc_matrix(A))

    def solve(self, b):
        return self._solve(np.asarray(b))


# ---------------------------------------------------------------------------------- AMG
def _set_coarse(h, Lc):
    """The coarse solve (L L^T)^-1 (gfem/linalg.py:126-133,178-181) as one explicit inverse,
    formed by LAPACK from the Cholesky factor."""
    import scipy.linalg as sla
    Lc = np.asarray(Lc, dtype=np.float64)
    Z = np.ascontiguousarray(sla.cho_solve((Lc, True), np.eye(Lc.shape[0])))
    L.check(L.load().exfem_amg_set_coarse_inverse(h, Z.shape[0], Z.ctypes.data))


class AMG:
    """SA-AMG (gfem/linalg.py:102-195): host setup, device V-cycle."""

    def __init__(self, A, X=None, max_coarse=300, degree=2, presmooth=1, rot_mask=None, log=None,
                 native=True):
        """A: scipy sparse (CSR, or BSR with 3x3 blocks as EbeOperator.assemble() returns it).
        native=False builds the levels with scipy's sparse products instead of the library's
        host SA level builder (the same hierarchy to rounding; kept as the reference route)."""
        t0 = time.time()
        n = A.shape[0]
        if X is not None and n == 3 * len(X):
            B, bs = rigid_body_modes(X), 3
        else:
            B, bs = np.ones((n, 1)), 1
        if native:
            self.levels_host, Ac = build_hierarchy_native(A, B, bs, max_coarse=max_coarse, log=log)
        else:
            A = sp.csr_matrix(A)
            self.levels_host, Ac = build_hierarchy(A, B, bs, max_coarse=max_coarse, log=log)
        self.Lc = coarse_cholesky(Ac.toarray())
        self.degree, self.presmooth = degree, presmooth
        self.nlevels = len(self.levels_host) + 1
        nnz = sum(lv["A"].nnz for lv in self.levels_host) + Ac.nnz
        self.complexity = float(nnz / max(A.nnz, 1))
        self.sizes = [lv["n"] for lv in self.levels_host] + [Ac.shape[0]]
        self.setup_time = time.time() - t0
        self.h = None
        self._op0 = None
        self._upload()

    @classmethod
    def from_operator(cls, op, X, max_coarse=300, max_levels=10, degree=2, presmooth=1, log=None,
                      keep_levels=False, nthreads=None, fp32=False):
        """The same hierarchy with its set-up on the device (exfem_sa_dev_*): level 0 is the
        operator's assembled FP64 matrix with its boundary mask (op: EbeOperator, after set_bc
        and set_ke), aggregation runs on host threads; the rigid-body candidates, strength graph,
        fit_candidates, smoothing, Galerkin products, transposes, diagonals, power estimates and
        the coarse Cholesky factor and inverse on the device, and the levels go to the V-cycle
        without leaving it. keep_levels: also copy the levels to
        the host (levels_host, for inspection). fp32: also make the FP32 copies of the cycle
        (set_precision(F32)) on the current stream."""
        from .amg_setup import _threads
        lib = L.load()
        device()
        t0 = time.time()
        self = cls.__new__(cls)
        self.h, self._op0 = None, op
        self.degree, self.presmooth = degree, presmooth
        n = 3 * len(X)
        if n != op.ndof:
            raise ValueError("X does not match the operator")
        Xh = np.ascontiguousarray(X, dtype=np.float64)
        cen = np.ascontiguousarray(Xh.mean(0))            # rigid_body_modes' centroid
        nt = nthreads or _threads()
        s = stream_ptr()
        sa = C.c_void_p()
        t_asm = time.time()
        L.check(lib.exfem_sa_dev_begin_rbm(C.byref(sa), op.h, Xh.ctypes.data, cen.ctypes.data, s))
        t_asm = time.time() - t_asm
        try:
            sizes, rhos, bs, cur = [n], [], 3, n
            while len(rhos) + 1 < max_levels and cur // bs > max_coarse:
                t1 = time.time()
                x0 = rho_start(cur)
                nagg, rho = C.c_int64(), C.c_double()
                L.check(lib.exfem_sa_dev_level(sa, 4.0 / 3.0, 1e-10, x0.ctypes.data, 15, nt,
                                               C.byref(nagg), C.byref(rho), s))
                if nagg.value <= 0 or nagg.value >= cur // bs:
                    break
                rhos.append(rho.value)
                if log:
                    log(f"SA level {len(rhos) - 1} (device): {cur} rows -> {nagg.value * 6} rows "
                        f"({time.time() - t1:.2f}s)")
                cur, bs = nagg.value * 6, 6
                sizes.append(cur)
            if not rhos:
                raise ValueError("matrix already at or below max_coarse: use a direct solve")
            t_lev = time.time()
            nl = len(rhos)
            nnz = []
            for lv in range(nl + 1):
                ab = C.c_int64()
                L.check(lib.exfem_sa_dev_info(sa, lv if lv < nl else -1, None, None, C.byref(ab), None,
                                              None, None))
                nnz.append(ab.value * (9 if lv == 0 else 36))
            tm = (C.c_double * 10)()
            L.check(lib.exfem_sa_dev_info(sa, -1, None, None, None, None, None, tm))
            self.setup_parts = dict(assemble_s=t_asm, graph_s=tm[1], aggregate_s=tm[2], fit_s=tm[3],
                                    products_s=tm[4], p_s=tm[5], ap_s=tm[6], pt_s=tm[7], ac_s=tm[8],
                                    rho_s=tm[9])
            h = C.c_void_p()
            L.check(lib.exfem_amg_create(C.byref(h), nl, degree))
            self.h = h
            t_to = time.time()
            L.check(lib.exfem_sa_dev_to_amg(sa, h, int(fp32), s))
            t_co = time.time()
            # coarse solve: Cholesky and inverse on the device; the eigenvalue route of a matrix
            # that is not positive definite (gfem/linalg.py:130-132) on the host
            ok = C.c_int()
            self.Lc = np.empty((cur, cur)) if keep_levels else None
            L.check(lib.exfem_sa_dev_coarse_solve(sa, h, None if self.Lc is None else self.Lc.ctypes.data,
                                                  C.byref(ok), s))
            if not ok.value:
                dense = np.empty((cur, cur))
                L.check(lib.exfem_sa_dev_coarse(sa, dense.ctypes.data, s))
                self.Lc = coarse_cholesky(dense)
                _set_coarse(h, self.Lc)
            self.setup_parts.update(levels_s=t_lev - t0 - t_asm, to_amg_s=t_co - t_to,
                                    coarse_s=time.time() - t_co)
            if keep_levels:
                self.levels_host = self._export_levels(sa, sizes, rhos)
        finally:
            lib.exfem_sa_dev_destroy(sa)
        self.nlevels = nl + 1
        self.sizes = sizes
        self.rhos = rhos
        self.complexity = float(sum(nnz) / max(nnz[0], 1))
        self._l0_matrix = False
        self.setup_time = time.time() - t0
        return self

    def _export_levels(self, sa, sizes, rhos):
        """Host copies of the device-built levels in build_hierarchy's format."""
        lib = L.load()
        s = stream_ptr()
        out = []
        for lv, rho in enumerate(rhos):
            n, nc = sizes[lv], sizes[lv + 1]
            bs = 3 if lv == 0 else 6
            mats = {}
            for which, key, (br, bc), nbr in ((0, "A", (bs, bs), n // bs), (1, "P", (bs, 6), n // bs),
                                              (2, "R", (6, bs), nc // 6)):
                if which == 0 and lv == 0:
                    mats[key] = self._op0.assemble(bc=True)
                    continue
                cnt = C.c_int64()
                if which == 0:
                    L.check(lib.exfem_sa_dev_info(sa, lv, None, None, C.byref(cnt), None, None, None))
                else:
                    L.check(lib.exfem_sa_dev_info(sa, lv, None, None, None, C.byref(cnt), None, None))
                m = max(cnt.value, 1)
                ptr, idx = np.empty(nbr + 1, np.int32), np.empty(m, np.int32)
                vals = np.empty((m, br, bc))
                L.check(lib.exfem_sa_dev_export(sa, lv, which, ptr.ctypes.data, idx.ctypes.data,
                                                vals.ctypes.data, s))
                mats[key] = sp.bsr_matrix((vals[:cnt.value], idx[:cnt.value], ptr),
                                          shape=(nbr * br, (nc // 6 if which == 1 else n // bs) * bc))
            agg = np.empty(n // bs, np.int32)
            L.check(lib.exfem_sa_dev_export(sa, lv, 3, None, agg.ctypes.data, None, s))
            d = mats["A"].diagonal()
            dinv = np.where(np.abs(d) > 0, 1.0 / np.where(d == 0, 1, d), 0.0)
            out.append(dict(A=mats["A"], P=mats["P"], R=mats["R"], dinv=dinv, rho=rho, n=n, nc=nc,
                            nagg=nc // 6, agg=agg))
        return out

    def _upload(self):
        """Device levels. Level 0's matrix stays on the host until a V-cycle without the EbE
        operator asks for it (the fused solves run level 0 on the operator)."""
        lib = L.load()
        device()
        nl = len(self.levels_host)
        if nl == 0:
            raise ValueError("matrix already at or below max_coarse: use a direct solve")
        h = C.c_void_p()
        L.check(lib.exfem_amg_create(C.byref(h), nl, self.degree))
        self.h = h
        self._l0_matrix = False
        for i, lv in enumerate(self.levels_host):
            Pp, Pi, Pv = csr_arrays(lv["P"])
            Rp, Ri, Rv = csr_arrays(lv["R"])
            dinv = np.ascontiguousarray(lv["dinv"], dtype=np.float64)
            Aa = None if i == 0 else csr_arrays(lv["A"])      # kept referenced across the call
            Ap, Ai, Av = (None, None, None) if Aa is None else (a.ctypes.data for a in Aa)
            L.check(lib.exfem_amg_set_level(h, i, lv["n"], lv["nc"], lv["rho"], dinv.ctypes.data,
                                            Ap, Ai, Av, Pp.ctypes.data, Pi.ctypes.data, Pv.ctypes.data,
                                            Rp.ctypes.data, Ri.ctypes.data, Rv.ctypes.data))
            # block-sparse copies for the batched cycle (level 0's A is the EbE operator there)
            for which, key in ((0, "A"), (1, "P"), (2, "R")):
                M = lv[key]
                if (which == 0 and i == 0) or not sp.isspmatrix_bsr(M):
                    continue
                br, bc = M.blocksize
                if (br, bc) not in ((3, 3), (3, 6), (6, 3), (6, 6)):
                    continue
                M.sort_indices()
                bp = np.ascontiguousarray(M.indptr, dtype=np.int32)
                bi = np.ascontiguousarray(M.indices, dtype=np.int32)
                bv = np.ascontiguousarray(M.data, dtype=np.float64)
                L.check(lib.exfem_amg_set_level_blocks(h, i, which, M.shape[0] // br, br, bc,
                                                       bp.ctypes.data, bi.ctypes.data, bv.ctypes.data))
        _set_coarse(h, self.Lc)

    def __del__(self):
        if getattr(self, "h", None) is not None and L is not None and L._lib is not None:
            try:
                L._lib.exfem_amg_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def set_precision(self, dtype):
        L.check(L.load().exfem_amg_set_precision(self.h, dtype))

    def info(self):
        return (f"AMG levels={self.nlevels} sizes={self.sizes} op.complexity={self.complexity:.2f}")

    def params(self):
        return self

    def _level0_matrix(self):
        if not self._l0_matrix and getattr(self, "_op0", None) is not None:
            L.check(L.load().exfem_amg_level0_from_op(self.h, self._op0.h, stream_ptr()))
            self._l0_matrix = True
        if not self._l0_matrix:
            Ap, Ai, Av = csr_arrays(self.levels_host[0]["A"])
            L.check(L.load().exfem_amg_set_level_matrix(self.h, 0, Ap.ctypes.data, Ai.ctypes.data,
                                                        Av.ctypes.data))
            self._l0_matrix = True

    def apply_multi(self, R, A0=None, dinv0=None, dtype=L.F64, cycle=None):
        """V(r) for every row of R (m, n), m <= 64 per batched cycle (the vmapped cycle of
        gfem/eigen.py:86-96); larger blocks run in slices of 64. cycle: precision of the cycle
        (default dtype; F32 on FP64 rows needs A0)."""
        if A0 is None:
            self._level0_matrix()
        Rd = as_dev(R, dtype)
        if Rd.ndim == 1:
            Rd = Rd[None]
        Z = torch().empty_like(Rd)
        op = None if A0 is None else A0.op
        cyc = dtype if cycle is None else cycle
        for i in range(0, Rd.shape[0], 64):
            Ri = Rd[i:i + 64].contiguous()
            Zi = Z[i:i + 64]
            L.check(L.load().exfem_amg_vcycle_multi(self.h, None if op is None else op.h,
                                                    vptr(dinv0), Ri.shape[0], vptr(Ri), vptr(Zi),
                                                    dtype, cyc, stream_ptr()))
        return Z

    def apply(self, r, A0=None, dinv0=None, dtype=L.F64):
        if A0 is None:
            self._level0_matrix()
        rd = as_dev(r, dtype)
        z = torch().empty_like(rd)
        op = None if A0 is None else A0.op
        L.check(L.load().exfem_amg_vcycle(self.h, None if op is None else op.h, vptr(dinv0),
                                          vptr(rd), vptr(z), dtype, stream_ptr()))
        return z

    def smooth(self, lvl, b, x=None, A0=None, dinv0=None, dtype=L.F64):
        """One call of the level smoother (cheb, gfem/linalg.py:161-175) for A x = b on level
        lvl from x (None: from x = 0); level 0 runs on A0 (a BcOperator) with dinv0 when given.
        Returns the new x."""
        if lvl == 0 and A0 is None:
            self._level0_matrix()
        bd = as_dev(b, dtype)
        xd = torch().empty_like(bd) if x is None else as_dev(x, dtype).clone()
        op = None if A0 is None else A0.op
        mf = 0 if A0 is None else int(A0.matrix_free)
        L.check(L.load().exfem_amg_smooth(self.h, None if op is None else op.h, vptr(dinv0), int(lvl),
                                          vptr(bd), vptr(xd), int(x is None), mf, dtype,
                                          stream_ptr()))
        return xd

    def make(self):
        """precond(params, r, A0=None, dinv0=None) -> V(r) (gfem/linalg.py:156-195)."""
        return lambda p, r, A0=None, dinv0=None: self.apply(
            r, A0, None if dinv0 is None else as_dev(dinv0))


__all__ = ["pcg", "AMG", "csr_to_device", "csr_mv", "rigid_body_modes", "Direct", "bc_matrix",
           "BcOperator", "Jacobi", "Chebyshev", "AMGPrecond", "precond_apply", "lmax_power"]
