"""Many load cases in one device solve (multi_load_cases, gfem/optimize.py:116-143)."""
import ctypes as C

import numpy as np

from . import _lib as L
from .model import KeBlock, as_dev, stream_ptr, torch, vptr
from .solvers import State, _Setup


def multi_load_cases(model, step, cases, cg_tol=1e-8, maxiter=20000, dtype=L.F64,
                     return_info=False):
    """Solve K u_i = f_i for every row of `cases` (ncase, ndof) with the BCs of `step`:
    block PCG, Jacobi preconditioner, each case with its own scalars and stopping test."""
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
    F = as_dev(cases, dtype)
    if F.ndim == 1:
        F = F[None]
    B = (F * free.to(F.dtype)[None, :]).contiguous()
    X = torch().empty_like(B)
    nr = B.shape[0]
    iters = (C.c_int * nr)()
    rel = (C.c_double * nr)()
    st = L.check(L.load().exfem_block_pcg_solve(op.h, nr, vptr(B), vptr(X), dtype, vptr(pinv),
                                                cg_tol, maxiter, iters, rel, stream_ptr()))
    if return_info:
        return X, dict(iters=np.array(iters[:]), relres=np.array(rel[:]), status=st)
    return X
