"""Many load cases in one device solve (multi_load_cases, gfem/optimize.py:116-143)."""
import ctypes as C

import numpy as np

from . import _lib as L
from .model import KeBlock, as_dev, stream_ptr, torch, vptr
from .solvers import State, _Setup


def multi_load_cases(model, step, cases, cg_tol=1e-8, maxiter=20000, dtype=L.F64,
                     return_info=False, precond="jacobi", amg=None):
    """Solve K u_i = f_i for every row of `cases` (ncase, ndof) with the BCs of `step`:
    block PCG, each case with its own scalars and stopping test. precond: "jacobi" (the
    reference's), "amg" (SA-AMG V-cycle batched over the cases, in dtype) or "amg32" (FP32
    V-cycle inside an FP64 solve). amg: a hierarchy to reuse (built from the assembled
    operator with the boundary mask otherwise)."""
    if precond not in ("jacobi", "amg", "amg32"):
        raise ValueError(f"unknown block preconditioner {precond!r}")
    if precond == "amg32" and dtype != L.F64:
        raise ValueError("amg32 is an FP64 solve with an FP32 cycle")
    m = model
    setup = _Setup(m, step, State(m))
    op = m.operator()
    for i, K in enumerate(m.element_stiffness(dtype=L.F64)):
        op.set_ke(i, K)
        if dtype == L.F32:
            op.set_ke(i, K.astype(L.F32))
    op.set_bc(setup.free_u8)
    free = as_dev(setup.free_np)
    d = op.diagonal()
    pinv = torch().where(free > 0, 1.0 / torch().where(d.abs() > 0, d, torch().ones_like(d)),
                         torch().ones_like(d))
    kind, h = L.PC_JACOBI, None
    if precond != "jacobi":
        from .linalg import AMG
        if amg is None:
            amg = AMG.from_operator(op, m.X)
        if precond == "amg32" or dtype == L.F32:
            amg.set_precision(L.F32)
        kind = L.PC_AMG_MIXED if precond == "amg32" else L.PC_AMG
        h = amg.h
    F = as_dev(cases, dtype)
    if F.ndim == 1:
        F = F[None]
    B = (F * free.to(F.dtype)[None, :]).contiguous()
    X = torch().empty_like(B)
    nr = B.shape[0]
    its, rels, st = [], [], 0
    for i in range(0, nr, 64):                  # at most 64 cases per block solve
        Bi, Xi = B[i:i + 64], X[i:i + 64]
        k = Bi.shape[0]
        iters = (C.c_int * k)()
        rel = (C.c_double * k)()
        st = max(st, L.check(L.load().exfem_block_pcg_solve_pc(
            op.h, k, vptr(Bi), vptr(Xi), dtype, kind, vptr(pinv), h, cg_tol, maxiter, iters, rel,
            stream_ptr())))
        its += iters[:]
        rels += rel[:]
    if return_info:
        return X, dict(iters=np.array(its), relres=np.array(rels), status=st, amg=amg)
    return X
