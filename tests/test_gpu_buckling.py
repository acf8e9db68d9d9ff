"""Linear buckling on the device (gfem/solvers.py:1074-1117, gfem/eigen.py:137-151): element
stresses (Model.stresses, gfem/model.py:129-136) and initial-stress matrices (geo_stiff,
gfem/model.py:113-118) against the oracle (FP64 1e-12, FP32 1e-6), and the load factors of a
clamped column by the device LOBPCG and the Lanczos branch against the oracle's dense
eigensolution of the same matrices (1e-6)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
E, NU = 200e3, 0.3


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("mat", ["linear", "neohookean"])
@pytest.mark.parametrize("etname", ["C3D8", "C3D8_FULL", "C3D10"])
def test_stress_and_geo_ke_match_oracle(gpu, etname, mat):
    import torch
    import exfem_amd as X
    from exfem_amd import _lib as L
    et = O.ELEMENTS[etname]
    if etname == "C3D10":
        Xn, c = O.box_hex(2.0, 1.0, 1.0, 2, 1, 1)
        Xn, c = O.hex_to_tet4(Xn, c)
        Xn, c = O.tet4_to_tet10(Xn, c)
    else:
        Xn, c = O.box_hex(3.0, 2.0, 2.0, 3, 2, 2)
    Xn = Xn + np.random.default_rng(41).uniform(-0.08, 0.08, Xn.shape)
    m = X.Model(Xn)
    mu, lam = O.lame(E, NU)
    if mat == "linear":
        m.add_block(c, X.ELEMENTS[etname], X.LinearElastic(E, NU))
        spec = ("linear", (E, NU))
    else:
        p = (mu / 2, 2 / (lam + 2 * mu / 3))
        m.add_block(c, X.ELEMENTS[etname], X.NeoHookean(*p))
        spec = ("neohookean", p)
    u = 0.03 * np.random.default_rng(42).standard_normal(3 * len(Xn))
    S = m.stresses(torch.as_tensor(u, device="cuda"))[0]["S"]
    S0 = O.element_cauchy(et, Xn[c], u.reshape(-1, 3)[c], spec)
    assert np.abs(_np(S) - S0).max() <= 1e-12 * np.abs(S0).max()
    K0 = O.geo_stiff(et, Xn[c], S0)
    op = m.operator()
    for storage in (L.KE_FULL, L.KE_SYM):
        K = _np(op.geo_ke(0, X.ELEMENTS[etname], S0, storage).to_array())
        assert np.abs(K - K0).max() <= 1e-12 * np.abs(K0).max()
    K32 = _np(op.geo_ke(0, X.ELEMENTS[etname], S0, L.KE_SYM, L.F32).to_array())
    assert np.abs(K32 - K0).max() <= 1e-6 * np.abs(K0).max()


def _column(n=3, L=20.0):
    Xn, c = O.box_hex(L, 1.0, 1.0, 10 * n, n, n)
    fix = np.nonzero(Xn[:, 0] < 1e-9)[0]
    tip = np.nonzero(Xn[:, 0] > L - 1e-9)[0]
    return Xn, c, fix, tip


def _oracle_factors(Xn, c, fix, tip, nmodes):
    nd = 3 * len(Xn)
    K0 = O.ops_assemble([(O.dofs(c), O.element_stiffness(O.HEX8, Xn[c], E, NU))], nd)
    free = np.ones(nd)
    free[(fix[:, None] * 3 + np.arange(3)).ravel()] = 0
    f = np.zeros(nd)
    f[tip * 3] = -1.0 / len(tip)
    u = spla.spsolve(O.bc_matrix(K0, free).tocsc(), free * f)
    sig = O.element_cauchy(O.HEX8, Xn[c], u.reshape(-1, 3)[c], ("linear", (E, NU)))
    KG = O.ops_assemble([(O.dofs(c), O.geo_stiff(O.HEX8, Xn[c], sig))], nd)
    lam, modes = O.buckling_factors(K0, KG, free, nmodes)
    return lam, K0, KG, free


@pytest.mark.parametrize("method", ["lanczos", "lobpcg"])
def test_column_buckling_matches_oracle(gpu, method):
    import exfem_amd as X
    Xn, c, fix, tip = _column()
    m = X.Model(Xn)
    m.add_block(c, X.HEX8, X.LinearElastic(E, NU))
    st = X.Step(kind="buckle", bcs=[(fix, d, 0.0) for d in range(3)], loads=[(tip, 0, -1.0 / len(tip))],
                nmodes=4, eigensolver=method)
    out = X.buckling(m, st)
    lam0, K0, KG, free = _oracle_factors(Xn, c, fix, tip, 4)
    lam = out["load_factors"]
    assert np.all(np.abs(lam[:4] / lam0 - 1) <= 1e-6), (lam, lam0)
    euler = np.pi ** 2 * E / 12 / (4 * 20.0 ** 2)
    assert 1.0 < lam[0] / euler < 1.2                    # the stiff C3D8 mesh, above Euler
    for k in range(4):                                   # each mode solves (K0 + lam KG) phi = 0
        phi = out["modes"][k] * free
        r = free * (K0 @ phi + lam[k] * (KG @ phi))
        assert np.linalg.norm(r) <= 1e-5 * np.linalg.norm(free * (K0 @ phi))
