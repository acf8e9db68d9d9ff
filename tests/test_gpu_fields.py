"""Thermal strain and eigenstrain fields on the device (Model._fields, gfem/model.py:446-463;
stress_gp :54-64; exfem_op_set_fields) against the oracle: internal forces of linear and J2
blocks and the J2 states (1e-12), the J2 tangent at the shifted strain (1e-6, the oracle's
central-difference tangent), the matrix-free product unchanged by the fields, and static solves
under a step temperature: free expansion == alpha dT X (1e-10) and a clamped bar with a
through-thickness gradient == the oracle's direct solve (1e-8)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
E, NU, ALPHA = 210e3, 0.3, 1.2e-5
TABLE = [[250.0, 0.0], [300.0, 0.01], [420.0, 0.3]]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _np(t):
    return t.detach().cpu().numpy()


def _assemble(conn, fe, ndof):
    f = np.zeros(ndof)
    d = (conn[:, :, None] * 3 + np.arange(3)).reshape(len(conn), -1)
    np.add.at(f, d.ravel(), fe.reshape(len(conn), -1).ravel())
    return f


def _mesh(kind):
    from exfem_amd import workloads
    if kind == "tet10":
        p = workloads.tet10_box(3, 3, 6, seed=1, E_range=(1.0, 1.0))
        return p.X * 10.0, p.conn
    p = workloads.cantilever(5, 3, 3, L=10.0, H=6.0, W=6.0, jitter=0.15, seed=2)
    return p.X, p.conn


def _fields(Xn, conn, seed):
    rng = np.random.default_rng(seed)
    T0 = 20.0 + rng.uniform(-5, 5, len(Xn))
    T = T0 + 100.0 * rng.uniform(0, 1, len(Xn))
    eig = 1e-3 * rng.standard_normal((len(conn), 9))
    return T, T0, eig


@pytest.mark.parametrize("kind", ["hex8", "hex8_plain", "tet10"])
@pytest.mark.parametrize("mat", ["linear", "j2"])
def test_field_forces_match_oracle(gpu, kind, mat):
    import torch
    import exfem_amd as X
    Xn, conn = _mesh(kind)
    et = {"hex8": X.HEX8, "hex8_plain": X.HEX8_PLAIN, "tet10": X.TET10}[kind]
    oet = {"hex8": O.HEX8, "hex8_plain": O.HEX8_PLAIN, "tet10": O.TET10}[kind]
    m = X.Model(Xn)
    material = X.LinearElastic(E, NU, alpha=ALPHA) if mat == "linear" else X.J2Plasticity(E, NU, TABLE, alpha=ALPHA)
    m.add_block(conn, et, material)
    T, T0, eig = _fields(Xn, conn, 11)
    m.T0 = T0
    fields = dict(T=torch.as_tensor(T, device="cuda"), eig=[torch.as_tensor(eig, device="cuda")])
    u = 1e-3 * np.random.default_rng(12).standard_normal(m.ndof)
    ud = torch.as_tensor(u, device="cuda")
    f, st = m.internal_force(ud, m.init_state(), fields)
    U = u.reshape(-1, 3)
    if mat == "linear":
        fe = O.element_force(oet, Xn[conn], U[conn], E, NU, Te=T[conn], T0e=T0[conn], alpha=ALPHA, eig=eig)
    else:
        om = O.J2Plasticity(E, NU, TABLE)
        om.alpha = ALPHA
        fe, so = O.element_force_state(oet, Xn[conn], U[conn], np.zeros((len(conn), oet.ngp, 19)), om,
                                       Te=T[conn], T0e=T0[conn], eig=eig)
        assert 0.05 < (so[..., 9] > 0).mean() < 0.98          # both branches
        assert rel(_np(st[0]), so) <= 1e-12
        Ko = O.element_stiffness_state(oet, Xn[conn], U[conn], np.zeros_like(so), om, Te=T[conn],
                                       T0e=T0[conn], eig=eig)
        K = _np(m.element_stiffness(ud, m.init_state(), fields)[0].to_array())
        assert rel(K, Ko) <= 1e-6
    f0 = _assemble(conn, fe, m.ndof)
    assert rel(_np(f), f0) <= 1e-12, rel(_np(f), f0)
    # without the fields the force is the plain one; the products never see the fields
    f_plain, _ = m.internal_force(ud, m.init_state())
    assert rel(_np(f_plain), f0) > 1e-3
    if mat == "linear":
        v = np.random.default_rng(13).standard_normal(m.ndof)
        y_f = _np(m.tangent_jvp(ud, None, torch.as_tensor(v, device="cuda"), fields))
        y_0 = _np(m.tangent_jvp(ud, None, torch.as_tensor(v, device="cuda")))
        assert rel(y_f, y_0) <= 1e-14


def test_free_thermal_expansion_and_clamped_gradient(gpu):
    import torch
    import exfem_amd as X
    Xn, conn = O.box_hex(6.0, 2.0, 2.0, 6, 2, 2)
    nn = len(Xn)
    dT = 80.0
    # free expansion on symmetry planes: u = alpha dT X exactly (linear elements)
    m = X.Model(Xn)
    m.add_block(conn, X.HEX8, X.LinearElastic(E, NU, alpha=ALPHA))
    m.T0 = np.full(nn, 20.0)
    bcs = [(np.nonzero(Xn[:, d] < 1e-9)[0], d, 0.0) for d in range(3)]
    st = X.Step(bcs=bcs, temperature=np.full(nn, 20.0 + dT))
    for solver in ("direct", "jacobi"):
        u, _, info = X.StaticSolver(m, verbose=False, linear_solver=solver, cg_tol=1e-12).solve(st)
        assert rel(_np(u), (ALPHA * dT * Xn).ravel()) <= 1e-10, solver
        assert np.abs(_np(info["reactions"])).max() <= 1e-8 * E * ALPHA * dT
    # clamped at x = 0, temperature rising through the thickness (bending): the oracle's
    # direct solve of K u = -f_int(0) on the free dofs
    m = X.Model(Xn)
    m.add_block(conn, X.HEX8, X.LinearElastic(E, NU, alpha=ALPHA))
    m.T0 = np.full(nn, 20.0)
    Tz = 20.0 + dT * Xn[:, 2] / 2.0
    fix = np.nonzero(Xn[:, 0] < 1e-9)[0]
    st = X.Step(bcs=[(fix, d, 0.0) for d in range(3)], temperature=Tz)
    u, _, _ = X.StaticSolver(m, verbose=False, linear_solver="direct").solve(st)
    nd = m.ndof
    free = np.ones(nd)
    free[(fix[:, None] * 3 + np.arange(3)).ravel()] = 0
    K = O.ops_assemble([(O.dofs(conn), O.element_stiffness(O.HEX8, Xn[conn], E, NU))], nd)
    f0 = _assemble(conn, O.element_force(O.HEX8, Xn[conn], np.zeros((len(conn), 8, 3)), E, NU,
                                         Te=Tz[conn], T0e=np.full(conn.shape, 20.0), alpha=ALPHA), nd)
    u0 = spla.spsolve(O.bc_matrix(K, free).tocsc(), -free * f0)
    assert np.abs(u0).max() > 1e-4
    assert rel(_np(u), u0) <= 1e-8
