"""MPC elimination on the device (Transformation, gfem/constraints.py:221-302; opA with T,
gfem/solvers.py:257-267) against the oracle restatement: the operator to 1e-12, static solves
of two meshes tied at a duplicated interface against the merged mesh (1e-8), a normal
constraint against the fixed dof it stands for."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
E, NU = 210e3, 0.3


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _tied(nx=4, ny=2, nz=2):
    """(merged X, conn), (split X, conn of two blocks, interface nodes of A and of B, paired)."""
    X, conn = O.box_hex(2.0 * nx, ny, nz, 2 * nx, ny, nz)
    Xa, ca = O.box_hex(1.0 * nx, ny, nz, nx, ny, nz)
    Xb, cb = O.box_hex(1.0 * nx, ny, nz, nx, ny, nz, origin=(nx, 0, 0))
    Xs = np.vstack([Xa, Xb])
    ia = np.nonzero(np.abs(Xa[:, 0] - nx) < 1e-9)[0]
    ib = np.nonzero(np.abs(Xb[:, 0] - nx) < 1e-9)[0] + len(Xa)
    ia = ia[np.lexsort((Xs[ia, 2], Xs[ia, 1]))]
    ib = ib[np.lexsort((Xs[ib, 2], Xs[ib, 1]))]
    return (X, conn), (Xs, ca, cb + len(Xa), ia, ib)


def _split_model(Xs, ca, cb, ia, ib):
    import exfem_amd as X
    m = X.Model(Xs)
    m.add_block(ca, X.HEX8, X.LinearElastic(E, NU))
    m.add_block(cb, X.HEX8, X.LinearElastic(E, NU))
    sd = (ib[:, None] * 3 + np.arange(3)).ravel()
    md = (ia[:, None] * 3 + np.arange(3)).ravel()[:, None]
    m.add_constraint(X.MPC(sd, md, np.ones((len(sd), 1)), "interface"))
    return m


def test_mpc_operator_matches_oracle(gpu):
    import exfem_amd as X
    from exfem_amd.solvers import State, _Setup
    _, (Xs, ca, cb, ia, ib) = _tied()
    m = _split_model(Xs, ca, cb, ia, ib)
    x0 = np.nonzero(Xs[:, 0] < 1e-9)[0]
    st = X.Step(bcs=[(x0, d, 0.0) for d in range(3)])
    setup = _Setup(m, st, State(m))
    op = m.operator()
    for i, K in enumerate(m.element_stiffness()):
        op.set_ke(i, K)
    op.set_bc(setup.free_u8)
    setup.tr.bind(op)
    ops = [(O.dofs(ca), O.element_stiffness(O.HEX8, Xs[ca], E, NU)),
           (O.dofs(cb), O.element_stiffness(O.HEX8, Xs[cb], E, NU))]
    nd = m.ndof
    T, smask = O.transformation(nd, [O.mpc_rows(m.constraints[0].slave, m.constraints[0].masters,
                                                m.constraints[0].weights)], setup.fixed)
    assert np.array_equal(smask, setup.tr.slave_mask)
    assert abs(T - setup.tr.T).max() == 0
    free = setup.free_np
    A = O.op_bc_mpc(lambda w: O.ops_matvec(ops, w, nd), free, T)
    v = np.random.default_rng(3).standard_normal(nd)
    assert rel(op.matvec(v, bc=True).cpu().numpy(), A(v)) <= 1e-12
    assert rel(op.matvec(v, bc=True, matrix_free=True).cpu().numpy(), A(v)) <= 1e-12
    assert rel(setup.tr.apply(v).cpu().numpy(), T @ v) <= 1e-15
    assert rel(setup.tr.apply_T(v).cpu().numpy(), T.T @ v) <= 1e-15


@pytest.mark.parametrize("solver", ["direct", "jacobi", "amg", "amg32", "mixed"])
def test_tied_meshes_solve_to_merged_mesh(gpu, solver):
    import exfem_amd as X
    (Xm, cm), (Xs, ca, cb, ia, ib) = _tied(10, 4, 4)     # above AMG's coarsest size
    m = _split_model(Xs, ca, cb, ia, ib)
    x0 = np.nonzero(Xs[:, 0] < 1e-9)[0]
    tip = np.nonzero(Xs[:, 0] > Xs[:, 0].max() - 1e-9)[0]
    st = X.Step(bcs=[(x0, d, 0.0) for d in range(3)], loads=[(tip, 2, -1.0)])
    s = X.StaticSolver(m, verbose=False, linear_solver=solver, cg_tol=1e-11)
    u, _, info = s.solve(st)
    u = u.cpu().numpy().reshape(-1, 3)
    # merged mesh by the oracle (direct)
    nd = Xm.shape[0] * 3
    ops = [(O.dofs(cm), O.element_stiffness(O.HEX8, Xm[cm], E, NU))]
    fixed = np.zeros(nd, bool)
    xm0 = np.nonzero(Xm[:, 0] < 1e-9)[0]
    fixed[(xm0[:, None] * 3 + np.arange(3)).ravel()] = True
    f = np.zeros(nd)
    f[np.nonzero(Xm[:, 0] > Xm[:, 0].max() - 1e-9)[0] * 3 + 2] = -1.0
    free = (~fixed).astype(float)
    ur = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, nd), free).tocsc(), free * f).reshape(-1, 3)
    key = lambda P: np.lexsort((P[:, 2], P[:, 1], P[:, 0]))
    keep = np.ones(len(Xs), bool)
    keep[ib] = False
    us = u[keep][key(Xs[keep])]
    assert rel(us, ur[key(Xm)]) <= 1e-8
    assert np.allclose(u[ib], u[ia], rtol=0, atol=1e-14 * np.abs(u).max())   # slaves follow


def test_normal_constraint_equals_fixed_dof(gpu):
    import exfem_amd as X
    Xm, cm = O.box_hex(8.0, 2.0, 2.0, 8, 2, 2)
    x0 = np.nonzero(Xm[:, 0] < 1e-9)[0]
    tip = np.nonzero(Xm[:, 0] > 8.0 - 1e-9)[0]
    res = []
    for use_mpc in (True, False):
        m = X.Model(Xm)
        m.add_block(cm, X.HEX8, X.LinearElastic(E, NU))
        bcs = [(x0, 1, 0.0), (x0, 2, 0.0)]
        if use_mpc:
            m.add_constraint(X.normal_constraint(x0, (3.0, 0.0, 0.0)))
        else:
            bcs.append((x0, 0, 0.0))
        st = X.Step(bcs=bcs, loads=[(tip, 2, -1.0)])
        u, _, _ = X.StaticSolver(m, verbose=False, linear_solver="jacobi", cg_tol=1e-11).solve(st)
        res.append(u.cpu().numpy())
    assert rel(res[0], res[1]) <= 1e-8
