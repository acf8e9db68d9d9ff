"""Springs, penalty contact and rigid walls on the device: their force / tangent kernels against
the oracle (gfem/discrete.py:44-120, gfem/contact.py:19-39,152-212,254-265), the resulting `rest`
operators inside every product of the operator (opA, gfem/solvers.py:257-267: plain, masked,
matrix-free, FP32, with MPC; the diagonal), and static solves with them.

Tolerances: kernels and products FP64 1e-12 (FP32 products 1e-6); solves 1e-8 against the
oracle's sparse direct solve; equilibrium to 1e-6 of the load."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
E, NU = 210000.0, 0.3


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("kind", ["axial", "bushing", "ground"])
def test_spring_kernels_match_oracle(gpu, kind):
    import torch
    from exfem_amd.discrete import GroundSpring, SpringBlock
    rng = np.random.default_rng(31)
    X = rng.uniform(0, 10, (40, 3))
    u = rng.standard_normal(120) * 0.5                     # large displacements: large rotations
    if kind == "ground":
        b = GroundSpring(rng.choice(40, 12, replace=False), [3.0, 5.0, 7.0])
    else:
        pairs = np.stack([rng.choice(40, 2, replace=False) for _ in range(15)])
        b = SpringBlock(pairs, 4.0 if kind == "axial" else [3.0, 5.0, 7.0], kind)
    fe, Ke = b.evaluate(torch.as_tensor(X, device="cuda"), u)
    for e in range(b.ne):
        f0, K0 = O.spring_element(kind, X[b.conn[e]], u.reshape(-1, 3)[b.conn[e]], b.k if kind != "axial" else 4.0)
        assert rel(_np(fe[e]), f0) <= 1e-12 and rel(_np(Ke[e]), K0) <= 1e-12, e


def test_contact_and_wall_kernels_match_oracle(gpu):
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.contact import RigidWall
    from exfem_amd.model import stream_ptr, vptr
    rng = np.random.default_rng(32)
    X = rng.uniform(0, 1, (60, 3))
    X[:20, 2] = 0.0
    u = 0.02 * rng.standard_normal(180)
    Xd, ud = torch.as_tensor(X, device="cuda"), torch.as_tensor(u, device="cuda")
    # selection from candidate facets
    ns, nc = 40, 5
    cand = np.stack([rng.choice(20, 3, replace=False) for _ in range(ns * nc)]).reshape(ns, nc, 3).astype(np.int32)
    slave = np.arange(20, 60, dtype=np.int32)
    nodes = torch.empty((ns, 4), dtype=torch.int32, device="cuda")
    act = torch.empty(ns, dtype=torch.float64, device="cuda")
    gap = torch.empty_like(act)
    dev = lambda a: torch.as_tensor(a, device="cuda")     # held by name: pointers outlive the call
    sd, cd = dev(slave), dev(cand)
    L.check(L.load().exfem_contact_select(ns, nc, vptr(Xd), vptr(ud), vptr(sd), vptr(cd), 0.6, 0.3, vptr(nodes),
                                          vptr(act), vptr(gap), stream_ptr()))
    n0, a0, g0 = O.contact_select(X, u, slave, cand, 0.6, 0.3)
    assert np.array_equal(_np(nodes), n0) and np.array_equal(_np(act), a0)
    assert np.abs(_np(gap) - g0).max() <= 1e-12 * np.abs(g0).max()
    # pairs: pressed (penetrating) and open, with multipliers
    k = rng.uniform(10, 100, ns)
    lam = np.where(rng.random(ns) < 0.3, 2.0, 0.0)
    act2 = np.ones(ns)
    act2[::7] = 0.0
    fe = torch.empty((ns, 12), dtype=torch.float64, device="cuda")
    Ke = torch.empty((ns, 12, 12), dtype=torch.float64, device="cuda")
    kd, ld, ad = dev(k), dev(lam), dev(act2)
    L.check(L.load().exfem_contact_pairs(ns, vptr(Xd), vptr(ud), vptr(nodes), vptr(kd), vptr(ld), vptr(ad),
                                         vptr(fe), vptr(Ke), None, stream_ptr()))
    x = (X + u.reshape(-1, 3))[n0]
    pressed = 0
    for p in range(ns):
        f0, K0 = O.contact_pair(x[p], k[p], lam[p], act2[p])
        pressed += bool(f0.any())
        assert np.abs(_np(fe[p]) - f0.ravel()).max() <= 1e-12 * max(np.abs(f0).max(), 1.0)
        assert np.abs(_np(Ke[p]) - K0).max() <= 1e-12 * max(np.abs(K0).max(), 1.0)
    assert 0 < pressed < ns
    w = RigidWall(np.arange(20), [0, 0, 0.001], [0, 0, 1.0], 50.0, motion=[0, 0, 0.01])
    pw = w.params(X, u, 0.5)
    fw, Kw = w.evaluate(pw)
    f0, K0 = O.wall_evaluate(X, u, np.arange(20), [0, 0, 0.006], [0, 0, 1.0], 50.0)
    assert np.abs(_np(fw) - f0).max() <= 1e-12 and np.abs(_np(Kw) - K0).max() <= 1e-12


def _rest_problem():
    """A jittered cantilever with bushings, ground springs and contact-pair matrices as rest."""
    from exfem_amd import workloads
    rng = np.random.default_rng(33)
    p = workloads.cantilever(8, 3, 3, jitter=0.1, seed=34)
    Ke = O.element_stiffness(O.HEX8, p.X[p.conn], E, NU)
    ops = [(O.dofs(p.conn), Ke)]
    gn = rng.choice(p.nn, 10, replace=False)
    g_ops = (O.dofs(gn[:, None]), np.stack([O.spring_element("ground", None, np.zeros((1, 3)), [1e4, 2e4, 3e4])[1]] * 10))
    sc = np.stack([rng.choice(p.nn, 2, replace=False) for _ in range(12)])
    s_K = np.stack([O.spring_element("axial", p.X[c], 0.01 * rng.standard_normal((2, 3)), 5e4)[1] for c in sc])
    s_ops = (O.dofs(sc), s_K)
    cn = np.stack([rng.choice(p.nn, 4, replace=False) for _ in range(6)])
    c_K = np.stack([O.contact_pair(np.vstack([p.X[c[1:]].mean(0) - [0, 0, 0.1], p.X[c[1:]]]), 1e5, 0.0, 1.0)[1]
                    for c in cn])
    c_ops = (O.dofs(cn), c_K)
    return p, ops, [g_ops, s_ops, c_ops]


def test_rest_operators_inside_every_product(gpu):
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.constraints import MPC, Transformation
    from exfem_amd.model import set_rest_ops
    p, ops, rest = _rest_problem()
    m = p.model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness(storage=L.KE_SYM)[0])
    op.set_ke(0, m.element_stiffness(storage=L.KE_SYM, dtype=L.F32)[0])
    set_rest_ops(op, [(d, torch.as_tensor(K, device="cuda")) for d, K in rest])
    allops = ops + rest
    x = np.random.default_rng(35).standard_normal(p.ndof)
    K_apply = lambda w: O.ops_matvec(allops, w, p.ndof)
    assert rel(_np(op.matvec(x)), K_apply(x)) <= 1e-12
    op.set_bc(p.free)
    free = p.free.astype(float)
    y0 = O.op_bc(K_apply, free)(x)
    assert rel(_np(op.matvec(x, bc=True)), y0) <= 1e-12
    assert rel(_np(op.matvec(x, bc=True, matrix_free=True)), y0) <= 1e-12
    assert rel(_np(op.matvec(x, bc=True, dtype=L.F32)).astype(float), y0) <= 1e-6
    y, d = op.matvec(x, bc=True, dot=True)
    assert abs(float(d) - x @ y0) <= 1e-12 * abs(x @ y0)
    diag = O.ops_diagonal(allops, p.ndof)
    assert rel(_np(op.diagonal()), diag) <= 1e-13
    # MPC: tie the x-displacements of two nodes
    tr = Transformation(p.ndof, [MPC([3 * 40], [[3 * 41]], [[1.0]])], ~p.free.astype(bool))
    fr = (p.free.astype(bool) & ~tr.slave_mask)
    op.set_bc(fr)
    tr.bind(op)
    y1 = O.op_bc_mpc(K_apply, fr.astype(float), tr.T)(x)
    assert rel(_np(op.matvec(x, bc=True)), y1) <= 1e-12
    Transformation(p.ndof, []).bind(op)
    # the batched product has no rest operators
    with pytest.raises(L.ExfemError):
        op.matvec_multi(np.stack([x, x]))
    # element vectors assembled per node
    fe = np.random.default_rng(36).standard_normal(rest[1][0].shape)
    yv = torch.zeros(p.ndof, dtype=torch.float64, device="cuda")
    op.rest_assemble(1, fe, yv)
    want = np.zeros(p.ndof)
    np.add.at(want, rest[1][0].ravel(), fe.ravel())
    assert rel(_np(yv), want) <= 1e-14
    # cleared rest: back to the element blocks alone
    set_rest_ops(op, [])
    op.set_bc(None)
    assert rel(_np(op.matvec(x)), O.ops_matvec(ops, x, p.ndof)) <= 1e-12


@pytest.mark.parametrize("solver", ["direct", "jacobi", "amg"])
def test_static_solve_on_springs_matches_direct(gpu, solver):
    """A jittered block held only by ground springs at x = 0 and joined to a second, offset
    block by bushings: StaticSolver equals the oracle's direct solve of the assembled
    element + spring stiffness (1e-8)."""
    import exfem_amd as X
    from exfem_amd import workloads
    a = workloads.cantilever(10, 4, 4, jitter=0.1, seed=37)
    b = workloads.cantilever(4, 4, 4, L=40.0, jitter=0.0, seed=0)
    Xb = b.X + [100.5, 0, 0]
    Xn = np.vstack([a.X, Xb])
    m = X.Model(Xn)
    m.add_block(np.vstack([a.conn, b.conn + a.nn]), X.HEX8, X.LinearElastic(E, NU))
    x0 = np.nonzero(a.X[:, 0] < 1e-9)[0]
    m.add_ground_spring(x0, [5e4, 5e4, 5e4])
    ia = np.nonzero(np.abs(a.X[:, 0] - 100.0) < 1e-9)[0]
    ib = np.nonzero(np.abs(Xb[:, 0] - 100.5) < 1e-9)[0] + a.nn
    key = lambda P: np.lexsort((P[:, 2], P[:, 1]))
    pairs = np.stack([ia[key(Xn[ia])], ib[key(Xn[ib])]], 1)
    m.add_spring(pairs, [2e5, 1e5, 1e5], kind="bushing")
    tip = np.nonzero(Xn[:, 0] > Xn[:, 0].max() - 1e-9)[0]
    st = X.Step(bcs=[], loads=[(tip, 2, -10.0)])
    u, _, info = X.StaticSolver(m, verbose=False, linear_solver=solver, cg_tol=1e-11).solve(st)
    nd = 3 * len(Xn)
    conn = np.vstack([a.conn, b.conn + a.nn])
    allops = [(O.dofs(conn), O.element_stiffness(O.HEX8, Xn[conn], E, NU)),
              (O.dofs(x0[:, None]), np.stack([np.diag([5e4] * 3)] * len(x0))),
              (O.dofs(pairs), np.stack([O.spring_element("bushing", None, np.zeros((2, 3)), [2e5, 1e5, 1e5])[1]] * len(pairs)))]
    f = np.zeros(nd)
    f[tip * 3 + 2] = -10.0
    ud = spla.spsolve(O.ops_assemble(allops, nd).tocsc(), f)
    assert rel(_np(u), ud) <= 1e-8, rel(_np(u), ud)


def _stack_on(k_contact):
    """Block A (top, 4x4x2) over block B (bottom, 4x4x2, clamped at z = 0), B's top face
    (triangulated) the master surface, A's bottom nodes the slaves, A's x/y held at its top."""
    from exfem_amd import mesh
    Xa, ca = mesh.box_hex(4.0, 4.0, 2.0, 4, 4, 2)
    Xb, cb = mesh.box_hex(4.0, 4.0, 2.0, 4, 4, 2)
    Xa = Xa + [0.0, 0.0, 2.0 - 1e-6]          # a first overlap: pressure needs a negative gap
    Xn = np.vstack([Xb, Xa])
    conn = np.vstack([cb, ca + len(Xb)])
    top_b = np.isclose(Xb[:, 2], 2.0)
    quads = cb[:, 4:8][top_b[cb[:, 4:8]].all(1)]
    tris = np.vstack([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]])
    slave = np.nonzero(np.isclose(Xa[:, 2], 2.0))[0] + len(Xb)
    return Xn, conn, tris, slave, len(Xb)


def test_rigid_wall_and_penalty_contact_balance_the_load(gpu):
    """A bar on a rigid wall and a block pressed on another through penalty contact: every
    slave closes, the contact forces balance the applied load (1e-6), and the wall solve equals
    the direct solve with its active normal springs (1e-8)."""
    import torch
    import exfem_amd as X
    from exfem_amd import mesh
    from exfem_amd.contact import PenaltyContact, RigidWall
    # bar on a wall at z = 0: bottom nodes are the slaves; x, y fixed at the bottom corners
    Xn, conn = mesh.box_hex(2.0, 2.0, 6.0, 2, 2, 6)
    m = X.Model(Xn)
    m.add_block(conn, X.HEX8, X.LinearElastic(E, NU))
    bot = np.nonzero(Xn[:, 2] < 1e-9)[0]
    top = np.nonzero(Xn[:, 2] > 6.0 - 1e-9)[0]
    kw, z0 = 1e6, 1e-6                      # the wall starts 1e-6 inside the bar (active)
    m.add_interaction(RigidWall(bot, [0, 0, z0], [0, 0, 1], kw))
    P = -1000.0
    st = X.Step(bcs=[(bot, 0, 0.0), (bot, 1, 0.0)], loads=[(top, 2, P / len(top))], dt_init=0.5, dt_max=0.5)
    u, _, info = X.StaticSolver(m, verbose=False, linear_solver="direct", rtol=1e-10).solve(st)
    uz = _np(u).reshape(-1, 3)[bot, 2]
    assert (uz < z0).all()
    assert abs(kw * (z0 - uz).sum() - (-P)) <= 1e-6 * abs(P)
    nd = m.ndof
    ops = [(O.dofs(conn), O.element_stiffness(O.HEX8, Xn[conn], E, NU)),
           (O.dofs(bot[:, None]), np.stack([np.diag([0, 0, kw])] * len(bot)))]
    free = np.ones(nd)
    free[bot * 3] = free[bot * 3 + 1] = 0
    f = np.zeros(nd)
    f[top * 3 + 2] = P / len(top)
    f[bot * 3 + 2] += kw * z0                 # the wall's force at u = 0 (linear once all close)
    ud = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, nd), free).tocsc(), free * f)
    assert rel(_np(u), ud) <= 1e-8
    # two blocks in penalty contact
    Xn, conn, tris, slave, nb_nodes = _stack_on(1e6)
    m = X.Model(Xn)
    m.add_block(conn, X.HEX8, X.LinearElastic(E, NU))
    kc = 1e6
    c = PenaltyContact(Xn, tris, slave, kc, h=1.0)
    m.add_interaction(c)
    base = np.nonzero(Xn[:, 2] < 1e-9)[0]
    atop = np.nonzero(Xn[:, 2] > 4.0 - 2e-6)[0]
    st = X.Step(bcs=[(base, d, 0.0) for d in range(3)] + [(atop, 0, 0.0), (atop, 1, 0.0)],
                loads=[(atop, 2, P / len(atop))], dt_init=0.5, dt_max=0.5)
    s = X.StaticSolver(m, verbose=False, linear_solver="jacobi", rtol=1e-10, cg_tol=1e-12)
    u, _, info = s.solve(st)
    p = c.params(Xn, u)
    assert _np(p["active"]).all() and (_np(p["gap"]) < 0).all()
    R = _np(info["reactions"]).reshape(-1, 3)[base, 2].sum()
    assert abs(R - (-P)) <= 1e-6 * abs(P), R
    fe, _ = c.evaluate(p, stiff=False)
    assert abs(_np(fe).reshape(-1, 4, 3)[:, 0, 2].sum() - P) <= 1e-6 * abs(P)
