"""Hyperelastic blocks on the device (exfem_op_set_hyper; gfem/materials.py:124-227, finite
strain with F-bar, gfem/model.py:40-106) against the oracle: internal force (1e-12), element
tangents (FP64 1e-10 against the oracle's complex-step tangent, FP32 1e-6), the matrix-free
product K(u) v against the stored tangent (1e-11), and large-deflection Newton solves against the
oracle's own Newton solve (1e-8)."""
import functools

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
E, NU = 200e3, 0.3
KINDS = list(O.HYPER_KINDS)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _np(t):
    return t.detach().cpu().numpy()


def _params(kind):
    mu, lam = O.lame(E, NU)
    D = 2 / (lam + 2 * mu / 3)
    return dict(neohookean=(mu / 2, D), mooney_rivlin=(0.3 * mu, 0.2 * mu, D),
                yeoh=(mu / 2, -0.05 * mu, 0.01 * mu, D, 4 * D, 0.0),
                arruda_boyce=(mu, 2.5, D), gent=(mu, 8.0, D), svk=(mu, lam))[kind]


def _material(kind):
    import exfem_amd as X
    p = _params(kind)
    return dict(neohookean=lambda: X.NeoHookean(*p), mooney_rivlin=lambda: X.MooneyRivlin(*p),
                yeoh=lambda: X.Yeoh(*p), arruda_boyce=lambda: X.ArrudaBoyce(*p),
                gent=lambda: X.Gent(*p), svk=lambda: X.StVenantKirchhoff(E, NU))[kind]()


def _mesh(etname, seed=21):
    if etname == "C3D10":
        X, c = O.box_hex(2.0, 1.0, 1.0, 2, 1, 1)
        X, c = O.hex_to_tet4(X, c)
        X, c = O.tet4_to_tet10(X, c)
    else:
        X, c = O.box_hex(3.0, 2.0, 2.0, 3, 2, 2)
    X = X + np.random.default_rng(seed).uniform(-0.08, 0.08, X.shape)
    return X, c


def _model(etname, kind, X, c):
    import exfem_amd as XX
    m = XX.Model(X)
    m.add_block(c, XX.ELEMENTS[etname], _material(kind))
    return m


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("etname", ["C3D8", "C3D8_FULL", "C3D10"])
def test_hyper_force_and_tangent_match_oracle(gpu, etname, kind):
    import torch
    from exfem_amd import _lib as L
    et = O.ELEMENTS[etname]
    X, c = _mesh(etname)
    m = _model(etname, kind, X, c)
    nd = 3 * len(X)
    u = 0.05 * np.random.default_rng(22).standard_normal(nd)
    ud = torch.as_tensor(u, device="cuda")
    p = _params(kind)
    fe = O.hyper_element_force(et, X[c], u.reshape(-1, 3)[c], kind, p)
    f0 = np.zeros(nd)
    np.add.at(f0, O.dofs(c).ravel(), fe.reshape(len(c), -1).ravel())
    f, _ = m.internal_force(ud)
    assert rel(_np(f), f0) <= 1e-12, rel(_np(f), f0)
    K0 = O.hyper_element_stiffness(et, X[c], u.reshape(-1, 3)[c], kind, p)
    for storage in (L.KE_FULL, L.KE_SYM):
        K = _np(m.element_stiffness(ud, storage=storage)[0].to_array())
        assert np.abs(K - K0).max() <= 1e-10 * np.abs(K0).max(), (storage, np.abs(K - K0).max() / np.abs(K0).max())
    K32 = _np(m.element_stiffness(ud, storage=L.KE_SYM, dtype=L.F32)[0].to_array())
    assert np.abs(K32 - K0).max() <= 1e-6 * np.abs(K0).max()


@pytest.mark.parametrize("kind", ["neohookean", "mooney_rivlin", "gent"])
@pytest.mark.parametrize("etname", ["C3D8", "C3D10"])
def test_hyper_jvp_equals_stored_tangent(gpu, etname, kind):
    """Model.tangent_jvp (dual-number force derivative) == the stored tangent's product, plain
    and with the boundary mask."""
    import torch
    X, c = _mesh(etname, seed=23)
    m = _model(etname, kind, X, c)
    nd = 3 * len(X)
    rng = np.random.default_rng(24)
    u = torch.as_tensor(0.05 * rng.standard_normal(nd), device="cuda")
    v = rng.standard_normal(nd)
    y = _np(m.tangent_jvp(u, None, torch.as_tensor(v, device="cuda")))
    Ke = m.element_stiffness(u)
    y0 = _np(m.matvec(Ke, v))
    assert rel(y, y0) <= 1e-11, rel(y, y0)
    op = m.operator()
    free = np.ones(nd)
    free[:9] = 0
    op.set_bc(free)
    op.set_state(u, [None])
    yb = _np(op.matvec(v, bc=True, matrix_free=True))
    yb0 = free * O.ops_matvec([(O.dofs(c), _np(Ke[0].to_array()))], free * v, nd) + (1 - free) * v
    assert rel(yb, yb0) <= 1e-11
    op.set_bc(None)


def _oracle_newton(et, X, c, kind, p, fixed, f_ext, steps=4, tol=1e-12):
    """Load-stepped Newton on the oracle's force and complex-step tangent (sparse direct)."""
    nd = 3 * len(X)
    free = np.ones(nd)
    free[fixed] = 0
    u = np.zeros(nd)
    for s in range(1, steps + 1):
        fe_ = f_ext * s / steps
        for _ in range(30):
            ue = u.reshape(-1, 3)[c]
            fe = O.hyper_element_force(et, X[c], ue, kind, p)
            fint = np.zeros(nd)
            np.add.at(fint, O.dofs(c).ravel(), fe.reshape(len(c), -1).ravel())
            r = free * (fint - fe_)
            if np.linalg.norm(r) <= tol * np.linalg.norm(fe_):
                break
            K = O.ops_assemble([(O.dofs(c), O.hyper_element_stiffness(et, X[c], ue, kind, p))], nd)
            u = u - spla.spsolve(O.bc_matrix(K, free).tocsc(), r)
    return u


P_TIP = -150.0


@functools.lru_cache(maxsize=1)
def _cantilever_oracle():
    Xn, c = O.box_hex(10.0, 1.0, 1.0, 10, 2, 2)
    fix = np.nonzero(Xn[:, 0] < 1e-9)[0]
    tip = np.nonzero(Xn[:, 0] > 10.0 - 1e-9)[0]
    f = np.zeros(3 * len(Xn))
    f[tip * 3 + 2] = P_TIP / len(tip)
    fixed = (fix[:, None] * 3 + np.arange(3)).ravel()
    return _oracle_newton(O.HEX8, Xn, c, "neohookean", _params("neohookean"), fixed, f)


@pytest.mark.parametrize("solver,tangent", [("direct", "ebe"), ("jacobi", "jvp"), ("cheby", "ebe")])
def test_hyper_cantilever_newton_matches_oracle(gpu, solver, tangent):
    """A NeoHookean C3D8 cantilever bent to a tip deflection of about a fifth of its length:
    StaticSolver's equilibrium == the oracle's Newton solution (1e-8), reactions balance the
    load."""
    import exfem_amd as X
    Xn, c = O.box_hex(10.0, 1.0, 1.0, 10, 2, 2)
    kind = "neohookean"
    p = _params(kind)
    m = X.Model(Xn)
    m.add_block(c, X.HEX8, X.NeoHookean(*p))
    fix = np.nonzero(Xn[:, 0] < 1e-9)[0]
    tip = np.nonzero(Xn[:, 0] > 10.0 - 1e-9)[0]
    P = P_TIP
    st = X.Step(bcs=[(fix, d, 0.0) for d in range(3)], loads=[(tip, 2, P / len(tip))],
                dt_init=0.25, dt_max=0.25)
    s = X.StaticSolver(m, verbose=False, linear_solver=solver, tangent=tangent, rtol=1e-10,
                       cg_tol=1e-12)
    u, _, info = s.solve(st)
    u = _np(u)
    u0 = _cantilever_oracle()
    assert abs(u0.reshape(-1, 3)[tip, 2].mean()) > 1.5          # large deflection
    assert rel(u, u0) <= 1e-8, rel(u, u0)
    R = _np(info["reactions"]).reshape(-1, 3)[fix, 2].sum()
    assert abs(R + P) <= 1e-6 * abs(P)
