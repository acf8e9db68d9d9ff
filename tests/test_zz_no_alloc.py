"""exfem_pcg_solve_noalloc (PCG with its vectors in the library's static device array) against
exfem_pcg_solve (the workspace form), through the C ABI: bitwise the same solution, iterations,
relres and status, on two calls in a row with the second on a smaller model, for the Jacobi,
Chebyshev and AMG preconditioners (FP64) and FP32 Jacobi; a model beyond the array's limit is an
error before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LIMIT = 1397568


def _setup(p, dtype=0):
    import torch
    from exfem_amd import _lib as L
    m = p.model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness(storage=L.KE_SYM, dtype=dtype)[0])
    if dtype != L.F64:
        op.set_ke(0, m.element_stiffness(storage=L.KE_SYM)[0])
    op.set_bc(p.free)
    d = op.diagonal()
    free = torch.as_tensor(p.free.astype(float), device="cuda")
    pinv = torch.where(free > 0, 1.0 / torch.where(d.abs() > 0, d, torch.ones_like(d)), torch.ones_like(d))
    tdt = torch.float64 if dtype == L.F64 else torch.float32
    b = (free * torch.as_tensor(p.f, device="cuda")).to(tdt)
    return m, op, pinv, b


def _both(op, b, pinv, pc, lmax=0.0, amg=None, dtype=0, tol=1e-10):
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import _workspace
    from exfem_amd.model import stream_ptr, vptr
    lib = L.load()
    out = []
    for noalloc in (False, True):
        x = torch.full_like(b, float("nan"))
        it, rr = C.c_int(), C.c_double()
        if noalloc:
            st = lib.exfem_pcg_solve_noalloc(op.h, vptr(b), vptr(x), dtype, pc, vptr(pinv), lmax,
                                             amg, 0, tol, 20000, C.byref(it), C.byref(rr),
                                             stream_ptr())
        else:
            ws = _workspace(op.ndof)
            st = lib.exfem_pcg_solve(ws.h, op.h, vptr(b), vptr(x), dtype, pc, vptr(pinv), lmax,
                                     amg, 0, tol, 20000, C.byref(it), C.byref(rr), stream_ptr())
        torch.cuda.synchronize()
        assert st >= 0, lib.exfem_last_error()
        out.append((x.clone(), it.value, rr.value, st))
    (x0, k0, r0, s0), (x1, k1, r1, s1) = out
    assert s0 == s1 == 0 and k0 == k1 and r0 == r1
    assert torch.equal(x0, x1)
    return k0


def test_noalloc_pcg_equals_workspace_pcg_on_two_calls(gpu):
    import torch
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    from exfem_amd.linalg import AMG, BcOperator, lmax_power
    from oracle import gfem_oracle as O
    big = workloads.cantilever(30, 8, 8, jitter=0.1, seed=21)        # 8,649 dof
    small = workloads.small_cantilever()                             # 1,575 dof
    m, op, pinv, b = _setup(big)
    k = _both(op, b, pinv, L.PC_JACOBI)
    assert k > 50
    lm = lmax_power(BcOperator(op), pinv)
    _both(op, b, pinv, L.PC_CHEBY, lmax=lm)
    ops = [(O.dofs(big.conn), O.element_stiffness(O.HEX8, big.X[big.conn], big.E, big.nu))]
    amg = AMG(O.bc_matrix(O.ops_assemble(ops, big.ndof), big.free.astype(float)), big.X)
    assert _both(op, b, pinv, L.PC_AMG, amg=amg.h) < k
    # the second call on a smaller model: the array keeps the larger call's contents
    m2, op2, pinv2, b2 = _setup(small)
    _both(op2, b2, pinv2, L.PC_JACOBI)
    m3, op3, pinv3, b3 = _setup(small, dtype=L.F32)
    _both(op3, b3, pinv3, L.PC_JACOBI, dtype=L.F32, tol=1e-5)


def test_noalloc_pcg_rejects_models_beyond_its_limit(gpu):
    import torch
    from exfem_amd import _lib as L
    from exfem_amd import mesh
    from exfem_amd.materials import LinearElastic
    from exfem_amd.model import Model, stream_ptr, vptr
    X, conn = mesh.box_hex(1.0, 1.0, 1.0, 77, 77, 77)               # 474,552 nodes
    m = Model(X)
    m.add_block(conn, "C3D8", LinearElastic(1.0, 0.3))
    op = m.operator()
    assert op.ndof > LIMIT
    b = torch.zeros(op.ndof, dtype=torch.float64, device="cuda")
    x = torch.empty_like(b)
    it, rr = C.c_int(), C.c_double()
    st = L.load().exfem_pcg_solve_noalloc(op.h, vptr(b), vptr(x), L.F64, L.PC_NONE, None, 0.0,
                                          None, 0, 1e-8, 10, C.byref(it), C.byref(rr), stream_ptr())
    assert st == -1 and b"scratch limit" in L.load().exfem_last_error()
