"""Parity of the HIP path (through the C ABI) with the CPU oracle and the golden fixtures.

Tolerances: FP64 element matrices and matvecs <= 1e-12 relative (norm-wise); FP32 storage
<= 1e-6; solutions <= 1e-8 relative displacement error vs the FP64 direct solve.
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E, NU = 210000.0, 0.3


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _np(t):
    return t.detach().cpu().numpy()


def _c0_model():
    import exfem_amd as X
    from exfem_amd import workloads
    p = workloads.small_cantilever()
    return p, p.model()


def _jittered(nx=12, ny=4, nz=4, seed=3):
    from exfem_amd import workloads
    return workloads.cantilever(nx, ny, nz, jitter=0.1, seed=seed)


# ------------------------------------------------------------------ element matrices
@pytest.mark.parametrize("storage", [0, 1])
@pytest.mark.parametrize("dtype", [0, 1])
def test_ke_hex8_matches_golden_and_oracle(gpu, storage, dtype):
    g = np.load(os.path.join(G, "c0_cantilever.npz"))
    p, m = _c0_model()
    K = m.element_stiffness(storage=storage, dtype=dtype)[0].numpy()
    tol = 1e-12 if dtype == 0 else 1e-6
    assert rel(K[g["Ke_sel"]], g["Ke"]) <= tol
    pj = _jittered()
    mj = pj.model()
    Kj = mj.element_stiffness(storage=storage, dtype=dtype)[0].numpy()
    Ko = O.element_stiffness(O.HEX8, pj.X[pj.conn], E, NU)
    assert rel(Kj, Ko) <= tol
    assert np.array_equal(Kj, np.swapaxes(Kj, 1, 2))


def test_ke_plain_and_tet10_match_oracle(gpu):
    import exfem_amd as X
    pj = _jittered()
    m = X.Model(pj.X)
    m.add_block(pj.conn, X.HEX8_PLAIN, X.LinearElastic(E, NU))
    K = m.element_stiffness()[0].numpy()
    assert rel(K, O.element_stiffness(O.HEX8_PLAIN, pj.X[pj.conn], E, NU)) <= 1e-12
    g = np.load(os.path.join(G, "tet10_small.npz"))
    m = X.Model(g["X"])
    m.add_block(g["conn"], X.TET10, X.LinearElastic(1.0, NU))
    for storage in (0, 1):
        Kt = m.element_stiffness(fields=dict(scale=[g["scale"]]), storage=storage)
        assert rel(Kt[0].numpy()[:6], g["Ke"]) <= 1e-12
        y = m.matvec(Kt, g["x"])
        assert rel(_np(y), g["y"]) <= 1e-12


def test_ke_pack_roundtrip(gpu):
    import exfem_amd as X
    from exfem_amd.model import KeBlock
    pj = _jittered(5, 2, 3)
    Ko = O.element_stiffness(O.HEX8, pj.X[pj.conn], E, NU)
    for storage in (0, 1):
        kb = KeBlock.from_array(X.HEX8, Ko, storage)
        assert np.array_equal(kb.numpy(), Ko if storage == 0 else np.triu(Ko) + np.swapaxes(np.triu(Ko, 1), 1, 2))


# ------------------------------------------------------------------ EbE matvec
@pytest.mark.parametrize("storage", [0, 1])
def test_matvec_matches_golden(gpu, storage):
    g = np.load(os.path.join(G, "c0_cantilever.npz"))
    p, m = _c0_model()
    Ks = m.element_stiffness(storage=storage)
    y = m.matvec(Ks, g["x"])
    assert rel(_np(y), g["y"]) <= 1e-12
    op = m.operator()
    op.set_bc(g["free"].astype(np.uint8))
    yb = op.matvec(g["x"], bc=True)
    assert rel(_np(yb), g["y_bc"]) <= 1e-12
    assert rel(_np(op.diagonal()), g["diag"]) <= 1e-13


def test_matvec_jittered_and_dot(gpu):
    import torch
    pj = _jittered(16, 6, 5, seed=7)
    m = pj.model()
    Ks = m.element_stiffness()
    ops = [(O.dofs(pj.conn), O.element_stiffness(O.HEX8, pj.X[pj.conn], E, NU))]
    x = np.random.default_rng(0).standard_normal(pj.ndof)
    y0 = O.ops_matvec(ops, x, pj.ndof)
    op = m.operator()
    op.set_ke(0, Ks[0])
    y, d = op.matvec(x, dot=True)
    assert rel(_np(y), y0) <= 1e-12
    assert abs(float(d) - x @ y0) <= 1e-12 * abs(x @ y0)


def test_matrix_free_and_internal_force_match_stored(gpu):
    pj = _jittered(10, 4, 3, seed=2)
    m = pj.model()
    Ks = m.element_stiffness()
    x = np.random.default_rng(3).standard_normal(pj.ndof)
    y = _np(m.matvec(Ks, x))
    ymf = _np(m.tangent_jvp(None, None, x))
    f = _np(m.internal_force(x)[0])
    f0 = O.internal_force(pj.X, [(O.HEX8, pj.conn, E, NU, None)], x, pj.nn)
    assert rel(ymf, y) <= 1e-12
    assert rel(f, f0) <= 1e-12


def test_matvec_is_bitwise_deterministic(gpu):
    import torch
    pj = _jittered(20, 6, 6, seed=1)
    m = pj.model()
    Ks = m.element_stiffness()
    x = np.random.default_rng(5).standard_normal(pj.ndof)
    op = m.operator()
    op.set_ke(0, Ks[0])
    ys = [_np(op.matvec(x)) for _ in range(3)]
    assert all(np.array_equal(ys[0], y) for y in ys[1:])


# ------------------------------------------------------------------ solvers
def test_pcg_jacobi_matches_oracle_and_golden(gpu):
    import exfem_amd as X
    from exfem_amd.linalg import BcOperator, Jacobi, pcg_device
    g = np.load(os.path.join(G, "c0_cantilever.npz"))
    p, m = _c0_model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness()[0])
    op.set_bc(p.free)
    free = p.free.astype(float)
    d = _np(op.diagonal())
    pinv = np.where(free > 0, 1 / np.where(np.abs(d) > 0, d, 1), 1)
    x, k, rr, st = pcg_device(BcOperator(op), free * p.f, Jacobi(pinv), tol=1e-10, maxiter=20000)
    assert st == 0 and rr <= 1e-10
    assert abs(k - int(g["pcg_iters"])) <= 1
    assert rel(_np(x), g["u"]) <= 1e-8
    assert rel(_np(x), g["pcg_x"]) <= 1e-8


def test_pcg_graph_replay_equals_eager_launches(gpu):
    """PCG iterations replayed from a captured HIP graph give bitwise the results of eager
    launches, for several right-hand sides in a row with eager work between the replays."""
    import torch
    from exfem_amd import linalg
    from exfem_amd.linalg import AMG, AMGPrecond, BcOperator, Jacobi, pcg_device
    from exfem_amd import _lib as L
    p, m = _c0_model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness()[0])
    op.set_ke(0, m.element_stiffness(dtype=L.F32)[0])
    op.set_bc(p.free)
    free = p.free.astype(float)
    d = _np(op.diagonal())
    pinv = np.where(free > 0, 1 / np.where(np.abs(d) > 0, d, 1), 1)
    ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], E, NU))]
    amg = AMG(O.bc_matrix(O.ops_assemble(ops, p.ndof), free), p.X)
    A = BcOperator(op)
    rng = np.random.default_rng(9)
    bs = [free * rng.standard_normal(p.ndof) for _ in range(3)]
    for M in (Jacobi(pinv), AMGPrecond(amg, A, pinv), AMGPrecond(amg, A, pinv, L.F64, cycle=L.F32)):
        linalg.pcg_graphs(False)
        ref = [pcg_device(A, b, M, tol=1e-10, maxiter=5000) for b in bs]
        linalg.pcg_graphs(True)
        for _ in range(2):
            for b, (x0, k0, r0, _) in zip(bs, ref):
                x, k, r, st = pcg_device(A, b, M, tol=1e-10, maxiter=5000)
                _ = op.matvec(b)                          # eager work between replays
                torch.cuda.synchronize()
                assert k == k0 and r == r0
                assert torch.equal(x, x0)


@pytest.mark.parametrize("solver", ["jacobi", "cheby", "amg", "amg32", "mixed"])
def test_static_solver_backends_agree_with_direct(gpu, solver):
    import exfem_amd as X
    g = np.load(os.path.join(G, "c0_cantilever.npz"))
    p, m = _c0_model()
    fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
    st = X.Step(bcs=[(fix, d, 0.0) for d in range(3)], dloads=[(p.f,)])
    s = X.StaticSolver(m, verbose=False, linear_solver=solver, cg_tol=1e-10)
    u, _, info = s.solve(st)
    u = _np(u)
    assert rel(u, g["u"]) <= 1e-8, (solver, rel(u, g["u"]))
    R = _np(info["reactions"]).reshape(-1, 3)[:, 2].sum()
    assert abs(R - 400.0) < 1e-6 * 400


def test_amg_vcycle_matches_oracle(gpu):
    import torch
    from exfem_amd.linalg import AMG, BcOperator
    p, m = _c0_model()
    ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], E, NU))]
    free = p.free.astype(float)
    A = O.bc_matrix(O.ops_assemble(ops, p.ndof), free)
    amg = AMG(A, p.X)
    op = m.operator()
    op.set_ke(0, m.element_stiffness()[0])
    op.set_bc(p.free)
    dinv0 = O.jacobi_inv(O.ops_diagonal(ops, p.ndof), free)
    r = free * np.random.default_rng(2).standard_normal(p.ndof)
    lv = [dict(A=l["A"], dinv=l["dinv"], rho=l["rho"], P=l["P"], R=l["R"]) for l in amg.levels_host]
    A0 = O.op_bc(lambda w: O.ops_matvec(ops, w, p.ndof), free)
    z0 = free * O.vcycle(lv, amg.Lc, free * r, A0=A0, dinv0=dinv0) + (1 - free) * r
    z = amg.apply(r, A0=BcOperator(op), dinv0=torch.as_tensor(dinv0, device="cuda"))
    assert rel(_np(z), z0) <= 1e-10
    # purely algebraic cycle (level 0 from the stored matrix)
    z1 = O.vcycle(lv, amg.Lc, r)
    assert rel(_np(amg.apply(r)), z1) <= 1e-10


@pytest.mark.parametrize("R", [1, 7, 64])
def test_amg_vcycle_multi_matches_single_cycles(gpu, R):
    """Batched V-cycle (the vmapped preconditioner of gfem/eigen.py:86-96) on R rows: every row
    within 1e-10 of the oracle's single V-cycle, level 0 on the EbE operator (batched product)
    and on the stored matrix; the FP32 cycle on FP64 rows - level 0 from the FP64 element
    matrices (assembled FP32 operator) - within 1e-5 of the FP64 one."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import AMG, BcOperator
    p, m = _c0_model()
    ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], E, NU))]
    free = p.free.astype(float)
    A = O.bc_matrix(O.ops_assemble(ops, p.ndof), free)
    amg = AMG(A, p.X)
    op = m.operator()
    op.set_ke(0, m.element_stiffness()[0])
    op.set_bc(p.free)
    dinv0 = O.jacobi_inv(O.ops_diagonal(ops, p.ndof), free)
    dd = torch.as_tensor(dinv0, device="cuda")
    Rm = free[None, :] * np.random.default_rng(R).standard_normal((R, p.ndof))
    lv = [dict(A=l["A"], dinv=l["dinv"], rho=l["rho"], P=l["P"], R=l["R"]) for l in amg.levels_host]
    A0 = O.op_bc(lambda w: O.ops_matvec(ops, w, p.ndof), free)
    Z = _np(amg.apply_multi(Rm, A0=BcOperator(op), dinv0=dd))
    Z1 = _np(amg.apply_multi(Rm))
    for k in range(R):
        r = Rm[k]
        z0 = free * O.vcycle(lv, amg.Lc, free * r, A0=A0, dinv0=dinv0) + (1 - free) * r
        assert rel(Z[k], z0) <= 1e-10, k
        assert rel(Z1[k], O.vcycle(lv, amg.Lc, r)) <= 1e-10, k
    amg.set_precision(L.F32)
    Z32 = _np(amg.apply_multi(Rm, A0=BcOperator(op), dinv0=dd, cycle=L.F32))
    for k in range(R):
        assert rel(Z32[k], Z[k]) <= 1e-5, k


def test_amg_pcg_matrix_free_level0(gpu):
    """AMG-PCG whose outer products and level-0 smoothing apply K matrix-free (FP64, and the
    FP32-inner mixed form) against a sparse direct solve of the oracle system, 1e-8."""
    import scipy.sparse.linalg as spla
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import AMG, AMGPrecond, BcOperator, pcg_device
    pj = _jittered(16, 4, 4, seed=9)
    m = pj.model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness(storage=1)[0])
    op.set_ke(0, m.element_stiffness(storage=1, dtype=L.F32)[0])
    op.set_bc(pj.free)
    ops = [(O.dofs(pj.conn), O.element_stiffness(O.HEX8, pj.X[pj.conn], E, NU))]
    free = pj.free.astype(float)
    A = O.bc_matrix(O.ops_assemble(ops, pj.ndof), free)
    b = free * pj.f
    x0 = spla.spsolve(A.tocsc(), b)
    amg = AMG(A, pj.X)
    d = O.ops_diagonal(ops, pj.ndof)
    pinv = torch.as_tensor(O.jacobi_inv(d, free), device="cuda")
    A64 = BcOperator(op, matrix_free=True)
    x, k, rr, st = pcg_device(A64, b, AMGPrecond(amg, A64, pinv), tol=1e-10, maxiter=500)
    assert st == 0 and rel(_np(x), x0) <= 1e-8
    A32 = BcOperator(op, matrix_free=True, dtype=L.F32)
    M32 = AMGPrecond(amg, A32, pinv, L.F32)
    bd = torch.as_tensor(b, device="cuda")
    xm = torch.zeros_like(bd)
    r = bd.clone()
    for _ in range(12):
        sc = float(torch.linalg.norm(r))
        dd, _, _, _ = pcg_device(A32, (r / sc).float(), M32, tol=2e-5, maxiter=500)
        xm = xm + dd.double() * sc
        r = bd - A64(xm)
        if float(torch.linalg.norm(r)) <= 1e-10 * float(torch.linalg.norm(bd)):
            break
    assert rel(_np(xm), x0) <= 1e-8


@pytest.mark.parametrize("mf", [False, True])
def test_amg_pcg_fp32_cycle_matches_direct(gpu, mf):
    """FP64 PCG preconditioned by the FP32 V-cycle (PC_AMG_MIXED) on a C3D10 box with E
    log-uniform in [1e3, 1e6], stored and matrix-free level 0: within 1e-8 of the direct FP64
    solve of the oracle system, in at most 10% (+2) more iterations than the FP64-cycle solve."""
    import torch
    import exfem_amd as X
    from exfem_amd import _lib as L, workloads
    from exfem_amd.linalg import AMG, AMGPrecond, BcOperator, pcg_device
    p = workloads.tet10_box(4, 4, 16, seed=0)
    m = X.Model(p.X)
    m.add_block(p.conn, X.TET10, X.LinearElastic(1.0, NU))
    fields = dict(scale=[p.scale])
    op = m.operator(fields)
    op.set_ke(0, m.element_stiffness(fields=fields, storage=L.KE_SYM, dtype=L.F64)[0])
    op.set_ke(0, m.element_stiffness(fields=fields, storage=L.KE_SYM, dtype=L.F32)[0])
    op.set_bc(p.free)
    ops = [(O.dofs(p.conn), O.element_stiffness(O.TET10, p.X[p.conn], 1.0, NU, p.scale))]
    free = p.free.astype(float)
    A = O.bc_matrix(O.ops_assemble(ops, p.ndof), free)
    b = free * p.f
    x0 = spla.spsolve(A.tocsc(), b)
    amg = AMG(A, p.X)
    pinv = torch.as_tensor(O.jacobi_inv(O.ops_diagonal(ops, p.ndof), free), device="cuda")
    A64 = BcOperator(op, matrix_free=mf)
    _, k64, _, st64 = pcg_device(A64, b, AMGPrecond(amg, A64, pinv), tol=1e-10, maxiter=3000)
    M = AMGPrecond(amg, A64, pinv, L.F64, cycle=L.F32)
    assert M.kind == L.PC_AMG_MIXED
    x, k, rr, st = pcg_device(A64, b, M, tol=1e-10, maxiter=3000)
    assert st64 == 0 and st == 0 and rr <= 1e-10
    assert rel(_np(x), x0) <= 1e-8
    assert k <= 1.1 * k64 + 2, (k, k64)


@pytest.mark.parametrize("storage", [0, 1])
@pytest.mark.parametrize("R", [1, 5, 64])
def test_multi_rhs_matvec_equals_single_matvecs(gpu, storage, R):
    import torch
    pj = _jittered(9, 4, 3, seed=6)
    m = pj.model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness(storage=storage)[0])
    op.set_bc(pj.free)
    U = torch.as_tensor(np.random.default_rng(R).standard_normal((R, pj.ndof)), device="cuda")
    for bc in (False, True):
        Y = op.matvec_multi(U, bc=bc)
        for r in range(R):
            y = op.matvec(U[r], bc=bc)
            assert float(torch.linalg.norm(Y[r] - y)) <= 1e-14 * float(torch.linalg.norm(y))


def test_multi_rhs_every_tile_shape(gpu):
    """Each autotuner candidate of the batched (MFMA) element pass gives the single-RHS result:
    C3D8 and C3D10, FP64 (1e-14) and FP32 (1e-6), R = 20 (a partial case tile)."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.workloads import tet10_box
    pj = _jittered(6, 4, 3, seed=3)
    mh = pj.model()
    pt = tet10_box(3, 3, 6, seed=1)
    mt = pt.model()
    for p, m, kw in ((pj, mh, {}), (pt, mt, dict(fields=dict(scale=[pt.scale])))):
        op = m.operator(kw.get("fields"))
        R = 20
        U = torch.as_tensor(np.random.default_rng(5).standard_normal((R, p.ndof)), device="cuda")
        for dt, tol in ((L.F64, 1e-14), (L.F32, 1e-6)):
            op.set_ke(0, m.element_stiffness(storage=1, dtype=dt, **kw)[0])
            op.set_bc(p.free)
            Ud = U if dt == L.F64 else U.float()
            ref = torch.stack([op.matvec(Ud[r], bc=True, dtype=dt) for r in range(R)])
            i = 0
            while True:
                try:
                    op.set_multi_shape(R, i, dtype=dt)
                except L.ExfemError:
                    break
                Y = op.matvec_multi(Ud, bc=True, dtype=dt)
                err = float(torch.linalg.norm((Y - ref).double()) / torch.linalg.norm(ref.double()))
                assert err <= tol, (p.name, dt, i, err)
                i += 1
            assert i >= 3
            op.set_multi_shape(R, None, dtype=dt)
            Y = op.matvec_multi(Ud, bc=True, dtype=dt)
            assert op.multi_shape(R, dtype=dt) is not None
            assert float(torch.linalg.norm((Y - ref).double()) / torch.linalg.norm(ref.double())) <= tol


@pytest.mark.parametrize("storage", [0, 1])
def test_device_assembly_matches_oracle(gpu, storage):
    """Device assembly == the oracle's ops_assemble, and with bc == its bc_matrix, on a model of
    two element blocks (the two halves of a jittered box) plus a node that no element touches."""
    import exfem_amd as X
    pj = _jittered(7, 3, 4, seed=8)
    Xn = np.vstack([pj.X, [[50.0, 50.0, 50.0]]])
    half = len(pj.conn) // 2
    m = X.Model(Xn)
    m.add_block(pj.conn[:half], X.HEX8, X.LinearElastic(E, NU))
    m.add_block(pj.conn[half:], X.HEX8, X.LinearElastic(E, NU))
    op = m.operator()
    Ks = m.element_stiffness(storage=storage)
    op.set_ke(0, Ks[0])
    op.set_ke(1, Ks[1])
    ndof = 3 * len(Xn)
    ops = [(O.dofs(pj.conn), O.element_stiffness(O.HEX8, pj.X[pj.conn], E, NU))]
    A0 = O.ops_assemble(ops, ndof)
    A = op.assemble().tocsr()
    assert abs(A - A0).max() <= 1e-12 * abs(A0).max()
    assert abs(A - A.T).max() <= 1e-12 * abs(A0).max()
    free = np.ones(ndof, np.uint8)
    free[:12] = 0
    free[-3:] = 0
    free[30] = 0
    op.set_bc(free)
    Ab = op.assemble(bc=True).tocsr()
    Ab0 = O.bc_matrix(A0, free.astype(float))
    assert abs(Ab - Ab0).max() <= 1e-12 * abs(A0).max()


def test_block_pcg_columns_equal_single_solves(gpu):
    import exfem_amd as X
    from exfem_amd.optimize import multi_load_cases
    pj = _jittered(10, 3, 3, seed=4)
    m = pj.model()
    fix = np.nonzero(pj.X[:, 0] < 1e-9)[0]
    step = X.Step(bcs=[(fix, d, 0.0) for d in range(3)])
    F = np.random.default_rng(0).standard_normal((5, pj.ndof))
    U, info = multi_load_cases(m, step, F, cg_tol=1e-10, return_info=True)
    ops = [(O.dofs(pj.conn), O.element_stiffness(O.HEX8, pj.X[pj.conn], E, NU))]
    U0 = O.multi_load_cases(ops, pj.ndof, pj.free.astype(float), F, cg_tol=1e-10)
    for k in range(5):
        assert rel(_np(U[k]), U0[k]) <= 1e-8
    assert (info["relres"] <= 1e-10).all()


@pytest.mark.parametrize("precond,dtype", [("amg", 0), ("amg32", 0), ("amg", 1)])
def test_block_pcg_amg_matches_oracle_and_single_solves(gpu, precond, dtype):
    """Block PCG with the batched V-cycle (AMG, FP32 cycle in an FP64 solve, all-FP32): every
    case within 1e-8 of a direct solve of the oracle system (all-FP32: 1e-3, the FP32 operator's
    own accuracy, about cond(K) x 6e-8 on this mesh), and an FP64 case takes the iterations of
    the single-vector AMG-PCG of the same system (+-1)."""
    import torch
    import exfem_amd as X
    from exfem_amd import _lib as L
    from exfem_amd.linalg import AMGPrecond, BcOperator, pcg_device
    from exfem_amd.optimize import multi_load_cases
    pj = _jittered(12, 4, 4, seed=4)
    m = pj.model()
    fix = np.nonzero(pj.X[:, 0] < 1e-9)[0]
    step = X.Step(bcs=[(fix, d, 0.0) for d in range(3)])
    F = np.random.default_rng(1).standard_normal((9, pj.ndof))
    tol = 1e-10 if dtype == 0 else 1e-6
    U, info = multi_load_cases(m, step, F, cg_tol=tol, return_info=True, precond=precond,
                               dtype=dtype)
    ops = [(O.dofs(pj.conn), O.element_stiffness(O.HEX8, pj.X[pj.conn], E, NU))]
    free = pj.free.astype(float)
    Ad = O.bc_matrix(O.ops_assemble(ops, pj.ndof), free).tocsc()
    for k in range(9):
        u0 = spla.spsolve(Ad, free * F[k])
        assert rel(_np(U[k]).astype(float), u0) <= (1e-8 if dtype == 0 else 1e-3), k
    assert info["status"] == 0 and (info["relres"] <= tol).all()
    assert info["iters"].max() <= 60, info["iters"]
    if dtype == 0:
        op = m.operator()
        op.set_ke(0, m.element_stiffness()[0])
        if precond == "amg32":                # the single FP32 cycle's level 0
            op.set_ke(0, m.element_stiffness(dtype=L.F32)[0])
        op.set_bc(pj.free)
        d = op.diagonal()
        fd = torch.as_tensor(free, device="cuda")
        pinv = torch.where(fd > 0, 1.0 / torch.where(d.abs() > 0, d, torch.ones_like(d)),
                           torch.ones_like(d))
        A = BcOperator(op)
        cyc = L.F32 if precond == "amg32" else None
        b = free * F[3]
        _, k1, _, st = pcg_device(A, b, AMGPrecond(info["amg"], A, pinv, L.F64, cycle=cyc),
                                  tol=tol, maxiter=500)
        assert st == 0 and abs(int(info["iters"][3]) - k1) <= 1, (info["iters"][3], k1)


# ------------------------------------------------------------------ full size (1.01M dof)
def test_full_size_matvec_properties(gpu):
    """200x40x40 jittered box: size-independent identities of K (linearity, symmetry, rigid-body
    null space away from the boundary mask), sym/full storage and matrix-free agreement."""
    import torch
    from exfem_amd import workloads
    p = workloads.jittered_cantilever()
    m = p.model()
    Kf = m.element_stiffness()
    Ks = m.element_stiffness(storage=1)
    op = m.operator()
    rng = np.random.default_rng(0)
    x = torch.as_tensor(rng.standard_normal(p.ndof), device="cuda")
    y = torch.as_tensor(rng.standard_normal(p.ndof), device="cuda")
    op.set_ke(0, Kf[0])
    Kx, Ky = op.matvec(x), op.matvec(y)
    Kxy = op.matvec(2.0 * x - 3.0 * y)
    assert float(torch.linalg.norm(Kxy - (2 * Kx - 3 * Ky)) / torch.linalg.norm(Kxy)) <= 1e-12
    assert abs(float(x @ Ky - y @ Kx)) <= 1e-12 * float(torch.linalg.norm(x) * torch.linalg.norm(Ky))
    t = torch.zeros(p.ndof, dtype=torch.float64, device="cuda")
    t[0::3] = 1.0
    assert float(torch.linalg.norm(op.matvec(t))) <= 1e-10 * float(torch.linalg.norm(Kx))
    op.set_ke(0, Ks[0])
    Kx_s = op.matvec(x)
    assert float(torch.linalg.norm(Kx_s - Kx) / torch.linalg.norm(Kx)) <= 1e-13
    Kx_mf = op.matvec(x, matrix_free=True)
    assert float(torch.linalg.norm(Kx_mf - Kx) / torch.linalg.norm(Kx)) <= 1e-12
    # with the boundary mask (the level-0 product of matrix-free AMG-PCG at this size)
    op.set_bc(p.free)
    Kx_b = op.matvec(x, bc=True)
    Kx_bmf = op.matvec(x, bc=True, matrix_free=True)
    assert float(torch.linalg.norm(Kx_bmf - Kx_b) / torch.linalg.norm(Kx_b)) <= 1e-12
    fixed = torch.as_tensor(~p.free.astype(bool), device="cuda")
    assert torch.equal(Kx_bmf[fixed], x[fixed])
    # sampled element rows against the oracle
    sel = np.r_[0:64, 159968:160032, 319936:320000]
    Ko = O.element_stiffness(O.HEX8, p.X[p.conn[sel]], E, NU)
    assert rel(Kf[0].numpy()[sel], Ko) <= 1e-12


def test_tet10_loguniform_E_mixed_precision_matches_direct(gpu):
    """C3D10 Delaunay box with per-element E log-uniform in [1e3, 1e6]: FP32 inner AMG-PCG in
    FP64 refinement reaches <= 1e-8 relative displacement error vs the FP64 direct solve."""
    import exfem_amd as X
    from exfem_amd import workloads
    p = workloads.tet10_box(4, 4, 16, seed=0)
    m = X.Model(p.X)
    m.add_block(p.conn, X.TET10, X.LinearElastic(1.0, NU))
    fix = np.nonzero(p.X[:, 2] < 1e-9)[0]
    st = X.Step(bcs=[(fix, d, 0.0) for d in range(3)], dloads=[(p.f,)])
    u, _, info = X.StaticSolver(m, verbose=False, linear_solver="mixed", cg_tol=1e-12).solve(st, scale=[p.scale])
    ops = [(O.dofs(p.conn), O.element_stiffness(O.TET10, p.X[p.conn], 1.0, NU, p.scale))]
    free = p.free.astype(float)
    u0 = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, p.ndof), free).tocsc(), free * p.f)
    assert rel(_np(u), u0) <= 1e-8


# ------------------------------------------------------------------ modal analysis
def test_lumped_mass_matches_oracle(gpu):
    """Device HRZ lumped mass (gfem/model.py:120-124, 501-510) == the oracle's, C3D8 and C3D10."""
    import exfem_amd as X
    from exfem_amd import workloads
    rho = 7.85e-9
    pj = _jittered(8, 3, 4, seed=11)
    m = X.Model(pj.X)
    m.add_block(pj.conn, X.HEX8, X.LinearElastic(E, NU, density=rho))
    M0 = O.lumped_mass(pj.X, [(O.HEX8, pj.conn, rho)])
    assert rel(_np(m.lumped_mass()), M0) <= 1e-13
    assert abs(m.total_mass() - M0.sum() / 3) <= 1e-12 * M0.sum()
    pt = workloads.tet10_box(3, 3, 6, seed=2)
    mt = X.Model(pt.X)
    mt.add_block(pt.conn, X.TET10, X.LinearElastic(1.0, NU, density=2.5))
    assert rel(_np(mt.lumped_mass()), O.lumped_mass(pt.X, [(O.TET10, pt.conn, 2.5)])) <= 1e-13


def test_modal_lobpcg_matches_lanczos_and_oracle(gpu):
    """modal with the device LOBPCG (batched EbE product, AMG V-cycle preconditioner) and with
    the Lanczos branch on the device-assembled matrix == the oracle's frequencies, C3D8
    cantilever 10 x 1 x 1 (40 x 4 x 4): Lanczos 1e-8, LOBPCG 1e-6 (the reference's
    tests/test_gpu_paths.py:17-25 tolerance between its two eigensolvers)."""
    import exfem_amd as X
    rho = 7.85e-9
    Xn, conn = O.box_hex(10.0, 1.0, 1.0, 40, 4, 4)
    m = X.Model(Xn)
    m.add_block(conn, X.HEX8, X.LinearElastic(E, NU, density=rho))
    fix = np.nonzero(Xn[:, 0] < 1e-9)[0]
    bcs = [(fix, d, 0.0) for d in range(3)]
    r_l = X.modal(m, X.Step(kind="frequency", bcs=bcs, nmodes=5, eigensolver="lanczos"))
    r_g = X.modal(m, X.Step(kind="frequency", bcs=bcs, nmodes=5, eigensolver="lobpcg"))
    ndof = 3 * len(Xn)
    free = np.ones(ndof)
    for d in range(3):
        free[fix * 3 + d] = 0
    ops = [(O.dofs(conn), O.element_stiffness(O.HEX8, Xn[conn], E, NU))]
    f0, _ = O.modal_frequencies(O.ops_assemble(ops, ndof), O.lumped_mass(Xn, [(O.HEX8, conn, rho)]), free, 5)
    assert np.allclose(r_l["freq"], f0, rtol=1e-8), (r_l["freq"], f0)
    assert np.allclose(r_g["freq"], f0, rtol=1e-6), (r_g["freq"], f0)
    assert np.isclose(r_g["total_mass"][0], rho * 10.0, rtol=1e-12)


def test_matvec_refuses_an_output_buffer_of_another_dtype_or_size(gpu):
    """The library writes as many values of the product's dtype as x holds: a smaller or
    differently typed output buffer is refused on the host, before any kernel runs."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    p = workloads.small_cantilever()
    m = p.model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness()[0])
    x = torch.ones(p.ndof, dtype=torch.float64, device="cuda")
    for y in (torch.empty(p.ndof, dtype=torch.float32, device="cuda"),
              torch.empty(p.ndof - 3, dtype=torch.float64, device="cuda")):
        with pytest.raises(ValueError):
            op.matvec(x, y)
        with pytest.raises(ValueError):
            op.matvec(x, y, matrix_free=True)
    y32 = torch.empty(p.ndof, dtype=torch.float32, device="cuda")
    op.matvec(x.float(), y32, matrix_free=True, dtype=L.F32)     # matching buffer: accepted
    assert torch.isfinite(y32).all()
