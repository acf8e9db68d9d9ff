"""Pins of the CPU oracle (oracle/gfem_oracle.py) against closed forms, invariants and
independent routes - the reference (JAX + PyAMG) cannot run here, so each restated function is
checked by something a plausible restatement mistake would fail."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

E, NU = 200e3, 0.3


@pytest.mark.parametrize("et", [O.HEX8, O.TET10])
def test_shape_derivatives_match_finite_differences(et):
    rng = np.random.default_rng(0)
    for p in list(et.gp) + [rng.uniform(0.05, 0.3, 3)]:
        h = 1e-6
        fd = np.stack([(et.Nfun(p + h * e) - et.Nfun(p - h * e)) / (2 * h) for e in np.eye(3)], 1)
        assert np.allclose(et.dNfun(p), fd, atol=1e-9)
        assert np.isclose(et.Nfun(p).sum(), 1.0)           # partition of unity
        assert np.allclose(et.dNfun(p).sum(0), 0.0, atol=1e-14)


def test_quadrature_integrates_polynomials():
    # 2x2x2 Gauss integrates xi^2 eta^2 zeta^2 exactly over [-1,1]^3: (2/3)^3
    g = O.HEX8.gp
    assert np.isclose((O.HEX8.gw * np.prod(g ** 2, 1)).sum(), (2 / 3) ** 3)
    # 4-point tet rule integrates quadratics: int x^2 over the unit tet = 1/60
    t = O.TET10.gp
    assert np.isclose((O.TET10.gw * t[:, 0] ** 2).sum(), 1 / 60)
    assert np.isclose(O.TET10.gw.sum(), 1 / 6)


def _brick(jitter=0.0, seed=0, shear=0.0):
    X, c = O.box_hex(1.0, 0.8, 1.2, 1, 1, 1)
    X = X.copy()
    X[:, 0] += shear * X[:, 2]
    if jitter:
        X += np.random.default_rng(seed).uniform(-jitter, jitter, X.shape)
    return X, c


def _tet10(seed=1):
    X, c = O.box_hex(1.0, 1.0, 1.0, 1, 1, 1)
    X, c = O.hex_to_tet4(X, c)
    X, c = O.tet4_to_tet10(X, c)
    return X, c


@pytest.mark.parametrize("kind", ["hex8", "hex8_plain", "tet10"])
def test_ke_symmetric_psd_with_rigid_body_nullspace(kind):
    if kind == "tet10":
        X, c = _tet10()
        et = O.TET10
    else:
        X, c = _brick(jitter=0.08)
        et = O.HEX8 if kind == "hex8" else O.HEX8_PLAIN
    K = O.element_stiffness(et, X[c], E, NU)
    for e in range(len(c)):
        Ke = K[e]
        assert np.allclose(Ke, Ke.T, atol=1e-10 * np.abs(Ke).max())
        B = O.rigid_body_modes(X[c[e]])
        assert np.abs(Ke @ B).max() < 1e-10 * np.abs(Ke).max()
        w = np.linalg.eigvalsh(Ke)
        assert (w > -1e-9 * w.max()).all()
        assert (w < 1e-9 * w.max()).sum() == 6


@pytest.mark.parametrize("kind", ["hex8", "hex8_plain", "tet10"])
def test_ke_is_derivative_of_force(kind):
    """Ke == d fe / du by central differences of the independently restated force routine."""
    if kind == "tet10":
        X, c = _tet10()
        et = O.TET10
        X, c = X, c[:2]
    else:
        X, c = _brick(jitter=0.1, seed=3)
        et = O.HEX8 if kind == "hex8" else O.HEX8_PLAIN
    Xe = X[c]
    sc = np.array([1.7] * len(c))
    K = O.element_stiffness(et, Xe, E, NU, sc)
    nn = et.nnode
    rng = np.random.default_rng(2)
    u = rng.standard_normal((len(c), nn, 3)) * 1e-3
    h = 1e-6
    for a in range(nn):
        for k in range(3):
            du = np.zeros_like(u)
            du[:, a, k] = h
            fd = (O.element_force(et, Xe, u + du, E, NU, sc) - O.element_force(et, Xe, u - du, E, NU, sc)) / (2 * h)
            assert np.allclose(K[:, :, 3 * a + k], fd.reshape(len(c), -1), rtol=1e-7,
                               atol=1e-7 * np.abs(K).max())
        # linear material: fe(u) == Ke u exactly (up to rounding)
    f = O.element_force(et, Xe, u, E, NU, sc).reshape(len(c), -1)
    assert np.allclose(f, np.einsum("eij,ej->ei", K, u.reshape(len(c), -1)), atol=1e-9 * np.abs(f).max())


def test_bbar_parallelepiped_equals_selective_reduced_integration():
    """On a parallelepiped the centre dilatation equals the volume average, so the reference
    B-bar element equals the textbook split: deviatoric part with the 2x2x2 rule + volumetric
    part with one point at the centre (bulk modulus kappa = lam + 2mu/3)."""
    X, c = _brick(shear=0.3)
    Xe = X[c]
    K = O.element_stiffness(O.HEX8, Xe, E, NU)[0]
    mu, lam = O.lame(E, NU)
    kappa = lam + 2 * mu / 3
    et = O.HEX8
    J = np.einsum("an,gai->gni", Xe[0], et.dN)
    b = np.einsum("gai,gij->gaj", et.dN, np.linalg.inv(J))
    v = np.linalg.det(J)
    J0 = Xe[0].T @ et.dN0
    b0 = et.dN0 @ np.linalg.inv(J0)
    V = np.linalg.det(J0) * 8.0
    assert np.isclose(v.sum(), V)
    nd = 24
    Kref = np.zeros((nd, nd))
    Id = np.eye(3)
    for g in range(8):
        for a in range(8):
            for bb in range(8):
                blk = mu * (Id * (b[g, a] @ b[g, bb]) + np.outer(b[g, bb], b[g, a])) \
                    - 2 * mu / 3 * np.outer(b[g, a], b[g, bb])
                Kref[3 * a:3 * a + 3, 3 * bb:3 * bb + 3] += v[g] * blk
    Kref += kappa * V * np.outer(b0.ravel(), b0.ravel())
    assert np.allclose(K, Kref, rtol=1e-12, atol=1e-12 * np.abs(K).max())
    # ... and differs from the plain element (volumetric locking relief)
    Kp = O.element_stiffness(O.HEX8_PLAIN, Xe, E, NU)[0]
    assert np.abs(K - Kp).max() > 1e-3 * np.abs(K).max()


def _uniaxial(et, n=2):
    X, c = O.box_hex(1, 1, 1, n, n, n)
    if et is O.TET10:
        X, c = O.hex_to_tet4(X, c)
        X, c = O.tet4_to_tet10(X, c)
    nd = 3 * len(X)
    fixed = np.zeros(nd, bool)
    fixed[np.nonzero(X[:, 0] < 1e-9)[0] * 3 + 0] = True
    fixed[np.nonzero(X[:, 1] < 1e-9)[0] * 3 + 1] = True
    fixed[np.nonzero(X[:, 2] < 1e-9)[0] * 3 + 2] = True
    right = np.nonzero(X[:, 0] > 1 - 1e-9)[0]
    fixed[right * 3] = True
    target = np.zeros(nd)
    target[right * 3] = 1e-3
    return X, c, fixed, target


@pytest.mark.parametrize("et", [O.HEX8, O.HEX8_PLAIN, O.TET10])
def test_patch_test_closed_form(et):
    """tests/test_verification.py:27-36 of the reference: uniaxial strain 1e-3 with free lateral
    faces gives sigma_xx = E 1e-3 and lateral contraction -nu 1e-3, exactly."""
    X, c, fixed, target = _uniaxial(et)
    nd = 3 * len(X)
    ops = [(O.dofs(c), O.element_stiffness(et, X[c], E, NU))]
    K = O.ops_assemble(ops, nd)
    free = (~fixed).astype(float)
    b = -free * (K @ target)
    A = O.bc_matrix(K, free)
    u = target + free * spla.spsolve(A.tocsc(), b)
    U = u.reshape(-1, 3)
    assert np.isclose(U[:, 1].min(), -NU * 1e-3, rtol=1e-6)
    # stress: sigma_xx from the element force of a single element equals E*1e-3 * area share
    f = O.internal_force(X, [(et, c, E, NU, None)], u, len(X))
    right = np.nonzero(X[:, 0] > 1 - 1e-9)[0]
    assert np.isclose(f.reshape(-1, 3)[right, 0].sum(), E * 1e-3, rtol=1e-9)


def test_matvec_forms_agree_and_match_assembled_csr():
    X, c = O.box_hex(3, 2, 2, 3, 2, 2)
    X = X + np.random.default_rng(0).uniform(-0.05, 0.05, X.shape) * (X[:, :1] > 0) * (X[:, :1] < 3)
    ops = [(O.dofs(c), O.element_stiffness(O.HEX8, X[c], E, NU))]
    nd = 3 * len(X)
    v = np.random.default_rng(1).standard_normal(nd)
    y1 = O.ops_matvec(ops, v, nd)
    y2 = O.gather_matvec(O.gather_plan([c], len(X)), ops, v, nd)
    y3 = O.ops_assemble(ops, nd) @ v
    y4 = O.internal_force(X, [(O.HEX8, c, E, NU, None)], v, len(X))
    for y in (y2, y3, y4):
        assert np.linalg.norm(y - y1) <= 1e-12 * np.linalg.norm(y1)


def test_uniform_mesh_stencil_closed_form():
    """Interior row of the assembled plain HEX8 matrix on a unit-cube mesh: for the constant-
    strain displacement u = (x, 0, 0) the interior force vanishes (row sums against linear
    fields), and K u = 0 for u = rigid translation."""
    X, c = O.box_hex(4, 4, 4, 4, 4, 4)
    ops = [(O.dofs(c), O.element_stiffness(O.HEX8_PLAIN, X[c], E, NU))]
    K = O.ops_assemble(ops, 3 * len(X))
    interior = np.nonzero(np.all((X > 0.5) & (X < 3.5), 1))[0]
    for field in (np.c_[X[:, 0], 0 * X[:, 0], 0 * X[:, 0]], np.c_[X[:, 1], X[:, 2], X[:, 0]]):
        r = (K @ field.ravel()).reshape(-1, 3)
        assert np.abs(r[interior]).max() < 1e-9 * E
    # diagonal of the 27-point stencil: 8 unit cubes x int[(lam+2mu) N_x^2 + mu (N_y^2+N_z^2)]
    # = 8 (lam + 4 mu) / 9 (each integral of a squared derivative is 1/9 on the unit cube)
    mu, lam = O.lame(E, NU)
    d = K.diagonal().reshape(-1, 3)[interior]
    assert np.allclose(d, 8 * (lam + 4 * mu) / 9, rtol=1e-13)
    # and every interior dof couples to the 27 nodes of its stencil (81 stored entries)
    nnz = np.diff(K.tocsr().indptr).reshape(-1, 3)[interior]
    assert (nnz == 81).all()


def test_pcg_matches_direct_and_counts_iterations():
    X, c = O.box_hex(4, 1, 1, 8, 2, 2)
    nd = 3 * len(X)
    ops = [(O.dofs(c), O.element_stiffness(O.HEX8, X[c], E, NU))]
    K = O.ops_assemble(ops, nd)
    free = np.ones(nd)
    free[np.nonzero(X[:, 0] < 1e-9)[0][:, None] * 3 + np.arange(3)] = 0
    b = free * np.random.default_rng(0).standard_normal(nd)
    A = O.op_bc(lambda w: O.ops_matvec(ops, w, nd), free)
    inv = O.jacobi_inv(O.ops_diagonal(ops, nd), free)
    x, k, rel = O.pcg(A, b, lambda r: inv * r, tol=1e-12, maxiter=5000)
    xd = spla.spsolve(O.bc_matrix(K, free).tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    assert rel <= 1e-12
    # stopping rule: CG on diag(1..m) with b = 1 converges in exactly m iterations
    m = 7
    D = np.arange(1.0, m + 1)
    x, k, rel = O.pcg(lambda v: D * v, np.ones(m), tol=1e-13, maxiter=100)
    assert k == m and np.allclose(x, 1 / D)
    # maxiter caps k and relres reports the recursive residual
    x, k, rel = O.pcg(lambda v: D * v, np.ones(m), tol=1e-13, maxiter=3)
    assert k == 3 and rel > 1e-3


def _similar_to_spectrum(lam, dinv, rng):
    """A = D^1/2 V diag(lam) V^T D^1/2 with D = diag(1/dinv): D^-1 A = diag(dinv) A has exactly
    the eigenvalues lam, with D^1/2-orthonormal eigenvectors. Returns (A, V)."""
    n = len(lam)
    V = np.linalg.qr(rng.standard_normal((n, n)))[0]
    Dh = np.diag(1.0 / np.sqrt(dinv))
    return Dh @ V @ np.diag(lam) @ V.T @ Dh, V


def _spectral(fun_of_lam, lam, V, dinv):
    """f(D^-1 A) = D^-1/2 V diag(f(lam)) V^T D^1/2 for the matrix of _similar_to_spectrum."""
    Dh = np.diag(1.0 / np.sqrt(dinv))
    Dhi = np.diag(np.sqrt(dinv))
    return Dhi @ V @ np.diag(fun_of_lam(lam)) @ V.T @ Dh


def _cheb_error_poly(k, lo, hi):
    """lam -> T_k((theta - lam)/delta) / T_k(theta/delta) on [lo, hi] (numpy's Chebyshev series,
    not a recurrence): the error polynomial of k steps of Chebyshev acceleration."""
    th, de = 0.5 * (hi + lo), 0.5 * (hi - lo)
    c = np.zeros(k + 1)
    c[k] = 1.0
    T = np.polynomial.chebyshev.chebval
    return lambda lam: T((th - np.asarray(lam)) / de, c) / T(th / de, c)


def test_chebyshev_smoother_error_propagator_is_t3():
    """cheb (gfem/linalg.py:161-175, deg 2) maps the error e = x - A^-1 b to
    p(D^-1 A) e with p = T_3((theta - lam)/delta) / T_3(theta/delta) on [rho/30, rho]: checked
    column by column on a matrix with a known spectrum. A smoother over [rho/10, rho] or of
    another degree gives a different propagator, which the same check rejects."""
    rng = np.random.default_rng(3)
    n = 16
    lam = np.concatenate([np.geomspace(0.004, 0.05, 5), np.linspace(0.06, 1.0, n - 5)])
    dinv = rng.uniform(0.5, 2.0, n)
    A, V = _similar_to_spectrum(lam, dinv, rng)
    rho = 1.1 * lam.max()
    Amv = lambda v: A @ v
    E = np.stack([O.cheb(Amv, dinv, rho, np.eye(n)[j], np.zeros(n), deg=2) for j in range(n)], 1)
    want = _spectral(_cheb_error_poly(3, rho / 30, rho), lam, V, dinv)
    scale = np.abs(want).max()
    assert np.abs(E - want).max() <= 1e-12 * scale
    # from x = 0 the smoother returns (I - p(D^-1 A)) A^-1 b
    b = rng.standard_normal(n)
    x = O.cheb(Amv, dinv, rho, np.zeros(n), b, deg=2)
    assert np.allclose(x, (np.eye(n) - want) @ np.linalg.solve(A, b), rtol=0, atol=1e-10 * np.abs(x).max())
    # the pin is discriminating: a wrong interval or degree is far outside the tolerance
    for wrong in (_cheb_error_poly(3, rho / 10, rho), _cheb_error_poly(2, rho / 30, rho),
                  _cheb_error_poly(4, rho / 30, rho), _cheb_error_poly(3, rho / 30, rho / 1.1)):
        assert np.abs(E - _spectral(wrong, lam, V, dinv)).max() > 1e-3 * scale
    # and the degree argument is honoured: deg = 4 gives T_5
    E4 = np.stack([O.cheb(Amv, dinv, rho, np.eye(n)[j], np.zeros(n), deg=4) for j in range(n)], 1)
    assert np.abs(E4 - _spectral(_cheb_error_poly(5, rho / 30, rho), lam, V, dinv)).max() <= 1e-12 * scale


def test_chebyshev_preconditioner_is_one_minus_t7():
    """cheby_precond (gfem/solvers.py:289-307, 6 steps from x = 0): M A = I - p(D^-1 A) with
    p = T_7((theta - lam)/delta) / T_7(theta/delta) on [lmax/30, lmax]."""
    rng = np.random.default_rng(4)
    n = 18
    lam = np.concatenate([np.geomspace(0.002, 0.03, 6), np.linspace(0.04, 2.0, n - 6)])
    dinv = rng.uniform(0.3, 3.0, n)
    A, V = _similar_to_spectrum(lam, dinv, rng)
    lmax = 1.1 * lam.max()
    M = O.cheby_precond(lambda v: A @ v, dinv, lmax)
    MA = np.stack([M(A[:, j]) for j in range(n)], 1)
    want = _spectral(lambda l: 1.0 - _cheb_error_poly(7, lmax / 30, lmax)(l), lam, V, dinv)
    scale = np.abs(want).max()
    assert np.abs(MA - want).max() <= 1e-11 * scale
    for wrong in (_cheb_error_poly(7, lmax / 10, lmax), _cheb_error_poly(6, lmax / 30, lmax),
                  _cheb_error_poly(8, lmax / 30, lmax)):
        assert np.abs(MA - _spectral(lambda l: 1.0 - wrong(l), lam, V, dinv)).max() > 1e-3 * scale
    # M is symmetric positive definite in the D^1/2 inner product (a valid PCG preconditioner)
    S = np.diag(1 / np.sqrt(dinv)) @ np.stack([M(np.eye(n)[j]) for j in range(n)], 1) @ np.diag(1 / np.sqrt(dinv))
    assert np.abs(S - S.T).max() <= 1e-12 * np.abs(S).max()
    assert np.linalg.eigvalsh(0.5 * (S + S.T)).min() > 0


def _power_closed_form(lam, V, dinv, v0, k):
    """|M^k v0| / |M^(k-1) v0| for M = D^-1 A of _similar_to_spectrum, from the eigenvalue
    powers (no iteration)."""
    Mk = lambda j: _spectral(lambda l: l ** j, lam, V, dinv) @ v0
    return np.linalg.norm(Mk(k)) / np.linalg.norm(Mk(k - 1))


def test_lmax_power_is_25_steps_from_the_sine_start():
    """_lmax (gfem/solvers.py:328-342): 1.1 x |M^25 v0| / |M^24 v0| with M = D^-1 A and
    v0 = sin(0.37 i + 0.1) free + 1e-3; a slowly converging spectrum (lam2/lam1 = 0.93) makes
    24 or 26 steps, or another start, differ well beyond rounding. With a separated spectrum it
    converges to 1.1 lam_max."""
    rng = np.random.default_rng(5)
    n = 14
    lam = np.concatenate([np.linspace(0.1, 0.8, n - 2), [0.93, 1.0]]) * 7.0
    dinv = rng.uniform(0.5, 2.0, n)
    A, V = _similar_to_spectrum(lam, dinv, rng)
    free = (rng.random(n) > 0.2).astype(float)
    v0 = np.sin(np.arange(n) * 0.37 + 0.1) * free + 1e-3
    got = O.lmax_power(lambda v: A @ v, dinv, free)
    want = 1.1 * _power_closed_form(lam, V, dinv, v0, 25)
    assert abs(got - want) <= 1e-12 * want
    for k in (24, 26):
        assert abs(got - 1.1 * _power_closed_form(lam, V, dinv, v0, k)) > 1e-6 * want
    other = np.cos(np.arange(n) * 0.37 + 0.1) * free + 1e-3
    assert abs(got - 1.1 * _power_closed_form(lam, V, dinv, other, 25)) > 1e-6 * want
    assert want <= 1.1 * lam.max() * (1 + 1e-12)
    lam2 = np.concatenate([np.linspace(0.1, 0.3, n - 1), [1.0]]) * 7.0
    A2, _ = _similar_to_spectrum(lam2, dinv, rng)
    assert abs(O.lmax_power(lambda v: A2 @ v, dinv, free) - 1.1 * 7.0) <= 1e-9 * 7.7


def test_rho_estimate_is_15_steps_from_the_seeded_normal_start():
    """AMG._rho (gfem/linalg.py:138-146): 1.1 x |M^15 x0| / |M^14 x0| with
    x0 = default_rng(0).standard_normal(n); 14 or 16 steps or another seed are rejected, and a
    separated spectrum gives 1.1 lam_max."""
    import scipy.sparse as sp
    rng = np.random.default_rng(6)
    n = 12
    lam = np.concatenate([np.linspace(0.05, 0.7, n - 2), [0.9, 1.0]]) * 3.0
    dinv = rng.uniform(0.5, 2.0, n)
    A, V = _similar_to_spectrum(lam, dinv, rng)
    x0 = np.random.default_rng(0).standard_normal(n)
    got = O.rho_estimate(sp.csr_matrix(A), dinv)
    want = 1.1 * _power_closed_form(lam, V, dinv, x0, 15)
    assert abs(got - want) <= 1e-12 * want
    for k in (14, 16):
        assert abs(got - 1.1 * _power_closed_form(lam, V, dinv, x0, k)) > 1e-6 * want
    x1 = np.random.default_rng(1).standard_normal(n)
    assert abs(got - 1.1 * _power_closed_form(lam, V, dinv, x1, 15)) > 1e-6 * want
    lam2 = np.concatenate([np.linspace(0.05, 0.2, n - 1), [1.0]]) * 3.0
    A2, _ = _similar_to_spectrum(lam2, dinv, rng)
    assert abs(O.rho_estimate(sp.csr_matrix(A2), dinv) - 3.3) <= 1e-9 * 3.3


def test_two_level_vcycle_exact_coarse_matches_dense():
    rng = np.random.default_rng(1)
    n, nc = 20, 5
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.linspace(1, 50, n)) @ Q.T
    P = rng.standard_normal((n, nc))
    import scipy.sparse as sp
    As = sp.csr_matrix(A)
    Ac = P.T @ A @ P
    Lc = O.coarse_factor(Ac)
    dinv = 1 / np.diag(A)
    lv = [dict(A=As, dinv=dinv, rho=O.rho_estimate(As, dinv), P=sp.csr_matrix(P), R=sp.csr_matrix(P.T))]
    b = rng.standard_normal(n)
    z = O.vcycle(lv, Lc, b)
    # dense: pre-smooth, exact coarse correction (with the 1e-13 shift), post-smooth
    sm = lambda x, bb: O.cheb(lambda v: A @ v, dinv, lv[0]["rho"], x, bb)
    x = sm(np.zeros(n), b)
    Acs = 0.5 * (Ac + Ac.T) + 1e-13 * np.abs(Ac).max() * np.eye(nc)
    x = x + P @ np.linalg.solve(Acs, P.T @ (b - A @ x))
    x = sm(x, b)
    assert np.allclose(z, x, rtol=1e-10, atol=1e-12)


def test_traction_weights_closed_form():
    """Config-0 tip face: 5x5 nodes of 5x5 quads of 5x5 area -> weights 6.25/12.5/25, total 400."""
    X, c = O.box_hex(100, 20, 20, 20, 4, 4)
    tip = X[:, 0] > 100 - 1e-6
    faces = c[:, [1, 2, 6, 5]]
    faces = faces[tip[faces].all(1)]
    w = O.quad4_weights(X, faces)
    f = O.traction_vector(w, (0, 0, -1.0)).reshape(-1, 3)
    assert np.isclose(f[:, 2].sum(), -400.0)
    ws = np.sort(np.unique(np.round(w[tip], 9)))
    assert np.allclose(ws, [6.25, 12.5, 25.0])


def test_cantilever_vs_timoshenko_and_reactions():
    """tests/test_verification.py:66-83 of the reference (TET10 beam 10x1x1, 3% of Timoshenko)."""
    L_, h = 10.0, 1.0
    X, c = O.box_hex(L_, h, h, 20, 2, 2)
    X, c = O.hex_to_tet4(X, c)
    X, c = O.tet4_to_tet10(X, c)
    nd = 3 * len(X)
    ops = [(O.dofs(c), O.element_stiffness(O.TET10, X[c], E, NU))]
    K = O.ops_assemble(ops, nd)
    fix = np.nonzero(X[:, 0] < 1e-9)[0]
    tip = np.nonzero(X[:, 0] > L_ - 1e-9)[0]
    free = np.ones(nd)
    for d in range(3):
        free[fix * 3 + d] = 0
    P = 1.0
    f = np.zeros(nd)
    f[tip * 3 + 1] = -P / len(tip)
    u = spla.spsolve(O.bc_matrix(K, free).tocsc(), free * f)
    I = h ** 4 / 12
    d_beam = P * L_ ** 3 / (3 * E * I) + P * L_ / (5 / 6 * E / (2 * (1 + NU)) * h * h)
    d = -u.reshape(-1, 3)[tip, 1].mean()
    assert abs(d - d_beam) / d_beam < 0.03
    R = ((K @ u - f) * (1 - free)).reshape(-1, 3)[:, 1].sum()
    assert abs(R - P) < 1e-6


def test_mixed_refinement_reaches_fp64_accuracy():
    X, c = O.box_hex(4, 1, 1, 8, 2, 2)
    nd = 3 * len(X)
    ops = [(O.dofs(c), O.element_stiffness(O.HEX8, X[c], E, NU))]
    K = O.ops_assemble(ops, nd)
    free = np.ones(nd)
    free[np.nonzero(X[:, 0] < 1e-9)[0][:, None] * 3 + np.arange(3)] = 0
    A = O.bc_matrix(K, free)
    A32 = A.astype(np.float32)
    b = free * np.random.default_rng(0).standard_normal(nd)
    inv32 = (1 / A.diagonal()).astype(np.float32)

    def inner(r32):
        x, k, _ = O.pcg(lambda v: A32 @ v, r32, lambda r: inv32 * r, tol=2e-5, maxiter=5000)
        return x, k

    x, its, steps = O.mixed_solve(lambda v: A @ v, inner, b, 1e-10)
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    assert steps > 1


def test_multi_load_cases_columns_equal_single_solves():
    X, c = O.box_hex(4, 1, 1, 6, 2, 2)
    nd = 3 * len(X)
    ops = [(O.dofs(c), O.element_stiffness(O.HEX8, X[c], E, NU))]
    free = np.ones(nd)
    free[np.nonzero(X[:, 0] < 1e-9)[0][:, None] * 3 + np.arange(3)] = 0
    F = np.random.default_rng(0).standard_normal((3, nd))
    U = O.multi_load_cases(ops, nd, free, F, cg_tol=1e-10)
    A = O.op_bc(lambda w: O.ops_matvec(ops, w, nd), free)
    inv = O.jacobi_inv(O.ops_diagonal(ops, nd), free)
    for k in range(3):
        x, _, _ = O.pcg(A, free * F[k], lambda r: inv * r, tol=1e-10, maxiter=20000)
        assert np.array_equal(x, U[k])


@pytest.mark.parametrize("etype", ["hex", "tet10"])
def test_c_gather_matvec_matches_numpy_oracle(etype):
    """oracle/gather_matvec.c (the CPU baseline) == the NumPy gather_matvec; thread count does
    not change the result."""
    from oracle import cpu
    from exfem_amd import workloads
    if etype == "hex":
        p = workloads.cantilever(7, 3, 3, jitter=0.1, seed=5)
        et = O.HEX8
        sc = None
    else:
        p = workloads.tet10_box(3, 3, 5, seed=3)
        et = O.TET10
        sc = p.scale
    Ke = O.element_stiffness(et, p.X[p.conn], p.E, p.nu, sc)
    x = np.random.default_rng(4).standard_normal(p.ndof)
    y0 = O.gather_matvec(O.gather_plan([p.conn], p.nn), [(O.dofs(p.conn), Ke)], x, p.ndof)
    y1 = cpu.GatherMatvec(p.conn, Ke, p.nn, threads=1)(x)
    y4 = cpu.GatherMatvec(p.conn, Ke, p.nn, threads=4)(x)
    assert np.abs(y1 - y0).max() <= 1e-14 * np.abs(y0).max()
    assert np.array_equal(y1, y4)


def test_c_jacobi_pcg_matches_numpy_oracle():
    """oracle_jacobi_pcg (the CPU baseline of solves/s) runs the reference's recurrence: on the
    1,575-dof cantilever and a jittered box, its iteration count equals the NumPy restatement's
    (+-1) and its solution matches the direct solve (1e-8)."""
    from oracle import cpu
    from exfem_amd import workloads
    for p in (workloads.small_cantilever(), workloads.cantilever(9, 3, 3, jitter=0.1, seed=6)):
        Ke = O.element_stiffness(O.HEX8, p.X[p.conn], p.E, p.nu)
        ops = [(O.dofs(p.conn), Ke)]
        free = p.free.astype(float)
        b = free * p.f
        A = O.op_bc(lambda w: O.ops_matvec(ops, w, p.ndof), free)
        inv = O.jacobi_inv(O.ops_diagonal(ops, p.ndof), free)
        x0, k0, r0 = O.pcg(A, b, lambda r: inv * r, tol=1e-10, maxiter=20000)
        x1, k1, r1 = cpu.GatherMatvec(p.conn, Ke, p.nn, threads=4).jacobi_pcg(b, free, tol=1e-10)
        assert abs(k1 - k0) <= 1 and r1 <= 1e-10
        xd = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, p.ndof), free).tocsc(), b)
        assert np.linalg.norm(x1 - xd) <= 1e-8 * np.linalg.norm(xd)


def test_lumped_mass_hrz_invariants():
    """HRZ lumping (gfem/model.py:120-124): each element's nodal masses add up to rho V_e; a
    parallelepiped HEX8 gets rho V / 8 on every node; on a straight-sided TET10 the corners share
    one value and the mid-edge nodes another, positive, with the ratio that the 4-point rule
    gives the closed-form shape functions (N_corner = L(2L-1), N_edge = 4 L_i L_j)."""
    rho = 7.85e-9
    X, c = _brick(shear=0.3)
    m = O.element_mass(O.HEX8, X[c], rho)
    V = O.geom(O.HEX8, X[c])[1].sum()
    assert np.isclose(V, 1.0 * 0.8 * 1.2)
    assert np.allclose(m, rho * V / 8, rtol=1e-13)
    Xj, cj = _brick(jitter=0.08, seed=4)
    mj = O.element_mass(O.HEX8, Xj[cj], rho)
    assert np.isclose(mj.sum(), rho * O.geom(O.HEX8, Xj[cj])[1].sum(), rtol=1e-13)
    # regular straight-sided TET10
    P = np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0], [0.5, np.sqrt(3) / 6, np.sqrt(2 / 3)]])
    X10 = np.vstack([P, [(P[i] + P[j]) / 2 for i, j in [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]]])
    mt = O.element_mass(O.TET10, X10[None], rho)[0]
    Vt = np.sqrt(2) / 12
    assert np.isclose(mt.sum(), rho * Vt, rtol=1e-13)
    assert np.all(mt > 0)
    assert np.allclose(mt[:4], mt[0], rtol=1e-12) and np.allclose(mt[4:], mt[4], rtol=1e-12)
    a, b = 0.5854101966249685, 0.1381966011250105
    corner = (a * (2 * a - 1)) ** 2 + 3 * (b * (2 * b - 1)) ** 2
    edge = 2 * (4 * a * b) ** 2 + 2 * (4 * b * b) ** 2
    assert np.isclose(mt[4] / mt[0], edge / corner, rtol=1e-12)


def test_modal_cantilever_matches_euler_bernoulli():
    """modal (gfem/solvers.py:1011-1046) on a C3D8 cantilever 10 x 1 x 1 (40 x 4 x 4): the first
    bending frequency is within 1% of 1.875104^2 / (2 pi L^2) sqrt(E I / (rho A)) (the reference's
    tests/test_advanced.py:112-125 closed form), the square section gives a double mode, and
    K phi = w^2 M phi holds for the returned modes."""
    rho = 7.85e-9
    X, conn = O.box_hex(10.0, 1.0, 1.0, 40, 4, 4)
    ops = [(O.dofs(conn), O.element_stiffness(O.HEX8, X[conn], E, NU))]
    K = O.ops_assemble(ops, 3 * len(X))
    M = O.lumped_mass(X, [(O.HEX8, conn, rho)])
    assert np.isclose(M.sum() / 3, rho * 10.0, rtol=1e-12)
    free = np.ones(3 * len(X))
    fix = np.nonzero(X[:, 0] < 1e-9)[0]
    for d in range(3):
        free[fix * 3 + d] = 0
    f, modes = O.modal_frequencies(K, M, free, 3)
    f1 = 1.875104 ** 2 / (2 * np.pi * 100.0) * np.sqrt(E * (1 / 12) / rho)
    assert abs(f[0] - f1) / f1 < 0.01
    assert np.isclose(f[0], f[1], rtol=1e-8)
    keep = free > 0
    Kf = K.tocsr()[keep][:, keep]
    for i in range(3):
        v = modes[i][keep]
        w2 = (2 * np.pi * f[i]) ** 2
        assert np.linalg.norm(Kf @ v - w2 * M[keep] * v) <= 1e-7 * np.linalg.norm(Kf @ v)


# ------------------------------------------------------------------ J2 plasticity
def _j2_uniaxial(mat, e1, s=None):
    """Stress of the return at eps = diag(e1, e2, e2) with e2 bisected so that sigma_22 = 0."""
    s = np.zeros(19) if s is None else s
    lo, hi = -abs(e1) - 1e-3, abs(e1) + 1e-3
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mat.ret(np.diag([e1, mid, mid]), s)[0][1, 1] > 0:
            hi = mid
        else:
            lo = mid
    return mat.ret(np.diag([e1, 0.5 * (lo + hi), 0.5 * (lo + hi)]), s)


@pytest.mark.parametrize("Hk", [0.0, 2000.0])
def test_j2_uniaxial_stress_closed_form(Hk):
    """Uniaxial stress with linear combined hardening: sigma = E (sy0 + H' eps) / (E + H'),
    H' = H_iso + Hk, above first yield; E eps below it. Elastic unloading from a plastic state
    follows slope E."""
    Ey, sy0, H = 210e3, 250.0, (400.0 - 250.0) / 0.2
    mat = O.J2Plasticity(Ey, 0.3, [[250, 0], [400, 0.2]], Hk=Hk)
    Hp = H + Hk
    for e1 in (5e-4, 2e-3, 6e-3, 3e-2):
        sig, s = _j2_uniaxial(mat, e1)
        cf = Ey * e1 if Ey * e1 <= sy0 else Ey * (sy0 + Hp * e1) / (Ey + Hp)
        assert abs(sig[0, 0] - cf) <= 1e-10 * cf
    sig_p, s_p = _j2_uniaxial(mat, 6e-3)
    sig_u, _ = _j2_uniaxial(mat, 5e-3, s_p)                 # unload by 1e-3 from the plastic state
    assert abs((sig_p[0, 0] - sig_u[0, 0]) - Ey * 1e-3) <= 1e-8 * Ey * 1e-3
    # perfect plasticity (one table row) and Johnson-Cook at T = 0
    pp = O.J2Plasticity(Ey, 0.3, [[300, 0]])
    assert abs(_j2_uniaxial(pp, 1e-2)[0][0, 0] - 300.0) <= 1e-9 * 300
    jc = O.J2Plasticity(Ey, 0.3, johnson_cook=dict(A=200.0, B=500.0, n=0.3))
    sig, s = _j2_uniaxial(jc, 1e-2)
    a = s[9]                                              # alpha = eps - sigma / E in uniaxial stress
    assert abs(a - (1e-2 - sig[0, 0] / Ey)) <= 1e-12
    assert abs(sig[0, 0] - (200.0 + 500.0 * (a + 1e-9) ** 0.3)) <= 1e-9 * sig[0, 0]


def test_j2_return_invariants():
    """After the return: yield consistency q(dev sigma - beta) == sy(alpha) at plastic points,
    trace(eps_p) == 0, backstress deviatoric; elastic points equal the linear material."""
    rng = np.random.default_rng(3)
    mat = O.J2Plasticity(210e3, 0.3, [[250, 0], [300, 0.01], [420, 0.3]], Hk=1500.0)
    lin = O.J2Plasticity(210e3, 0.3, [[1e12, 0]])
    nplastic = 0
    for _ in range(40):
        A = rng.standard_normal((3, 3)) * 4e-3
        eps = 0.5 * (A + A.T)
        s0 = np.zeros(19)
        sig, s1 = mat.ret(eps, s0)
        sig2, s2 = mat.ret(eps + 0.5 * eps, s1)           # a second step from the new state
        for sg, sn in ((sig, s1), (sig2, s2)):
            ep, a, beta = sn[:9].reshape(3, 3), sn[9], sn[10:].reshape(3, 3)
            assert abs(np.trace(ep)) <= 1e-15 and abs(np.trace(beta)) <= 1e-9
            xi = sg - np.trace(sg) / 3 * np.eye(3) - beta
            q = np.sqrt(1.5 * np.sum(xi * xi))
            if a > 0:
                nplastic += 1
                assert abs(q - mat.sy(a)) <= 1e-9 * mat.sy(a)
            else:
                assert q <= mat.sy(0.0) * (1 + 1e-12)
        mu, lam = O.lame(210e3, 0.3)
        small = 1e-4 * eps
        for m_, e_ in ((mat, small), (lin, eps)):
            lin_sig = lam * np.trace(e_) * np.eye(3) + 2 * mu * e_
            assert np.linalg.norm(m_.ret(e_, s0)[0] - lin_sig) <= 1e-13 * np.linalg.norm(lin_sig)
    assert nplastic >= 20


@pytest.mark.parametrize("et", [O.HEX8, O.HEX8_PLAIN, O.TET10])
def test_j2_element_tangent_is_the_force_derivative(et):
    """Ke of the state-carrying element (per-point tangent by central differences of the
    return) == d(fe)/du by central differences of element_force_state at a plastic state; with an
    unreachable yield stress both equal the linear element routines."""
    rng = np.random.default_rng(5)
    ref = O.box_hex(1.0, 1.0, 1.0, 1, 1, 1)
    X, conn = ref
    if et is O.TET10:
        X, conn = O.tet4_to_tet10(*O.hex_to_tet4(X, conn))
        conn = conn[:1]
    X = X + rng.uniform(-0.05, 0.05, X.shape)
    Xe = X[conn]
    ne, nn = Xe.shape[0], et.nnode
    mat = O.J2Plasticity(210e3, 0.3, [[250, 0], [400, 0.2]], Hk=800.0)
    ue = rng.standard_normal((ne, nn, 3)) * 4e-3
    se = np.zeros((ne, et.ngp, 19))
    _, s1 = O.element_force_state(et, Xe, ue, se, mat)    # a plastic history
    assert (s1[..., 9] > 0).any()
    ue2 = ue + rng.standard_normal((ne, nn, 3)) * 2e-3
    K = O.element_stiffness_state(et, Xe, ue2, s1, mat)
    h = 1e-8
    Kfd = np.zeros_like(K)
    for a in range(nn):
        for k in range(3):
            d = np.zeros_like(ue2)
            d[:, a, k] = h
            fp = O.element_force_state(et, Xe, ue2 + d, s1, mat)[0]
            fm = O.element_force_state(et, Xe, ue2 - d, s1, mat)[0]
            Kfd[:, :, 3 * a + k] = ((fp - fm) / (2 * h)).reshape(ne, -1)
    assert np.linalg.norm(K - Kfd) <= 1e-6 * np.linalg.norm(K)
    assert np.linalg.norm(K - np.swapaxes(K, 1, 2)) <= 1e-7 * np.linalg.norm(K)
    big = O.J2Plasticity(210e3, 0.3, [[1e12, 0]])
    Kl = O.element_stiffness(et, Xe, 210e3, 0.3)
    assert np.linalg.norm(O.element_stiffness_state(et, Xe, ue2, se, big) - Kl) <= 1e-6 * np.linalg.norm(Kl)
    fl = O.element_force(et, Xe, ue2, 210e3, 0.3)
    assert np.linalg.norm(O.element_force_state(et, Xe, ue2, se, big)[0] - fl) <= 1e-12 * np.linalg.norm(fl)


# ------------------------------------------------------------------ MPC transformation
def test_transformation_chains_cycles_and_dropped_slaves():
    """u_0 = u_1, u_1 = 0.5 u_5 resolves to u_0 = 0.5 u_5; a cycle u_2 = u_3, u_3 = u_2 leaves
    both slaves at zero; a prescribed slave is dropped; masters keep identity rows."""
    m1 = O.mpc_rows([0, 1], [[1], [5]], [[1.0], [0.5]])
    m2 = O.mpc_rows([2, 3, 4], [[3], [2], [6]], [[1.0], [1.0], [2.0]])
    fixed = np.zeros(8, bool)
    fixed[4] = True
    T, smask = O.transformation(8, [m1, m2], fixed)
    Td = T.toarray()
    assert np.array_equal(smask, [True, True, True, True, False, False, False, False])
    assert Td[0, 5] == 0.5 and Td[1, 5] == 0.5 and Td[0].sum() == 0.5
    assert not Td[2].any() and not Td[3].any()
    for i in (4, 5, 6, 7):
        assert Td[i, i] == 1.0 and Td[i].sum() == 1.0
    assert not Td[:, smask].any()


def _two_bars(nx=4, ny=2, nz=2):
    """A 2L x 1 x 1 bar as one mesh, and as two meshes whose interface nodes are duplicated."""
    X, conn = O.box_hex(2.0 * nx, ny, nz, 2 * nx, ny, nz)
    Xa, ca = O.box_hex(1.0 * nx, ny, nz, nx, ny, nz)
    Xb, cb = O.box_hex(1.0 * nx, ny, nz, nx, ny, nz, origin=(nx, 0, 0))
    Xs = np.vstack([Xa, Xb])
    cs = np.vstack([ca, cb + len(Xa)])
    ia = np.nonzero(np.abs(Xa[:, 0] - nx) < 1e-9)[0]
    ib = np.nonzero(np.abs(Xb[:, 0] - nx) < 1e-9)[0] + len(Xa)
    order_a = np.lexsort((Xs[ia, 2], Xs[ia, 1]))
    order_b = np.lexsort((Xs[ib, 2], Xs[ib, 1]))
    return (X, conn), (Xs, cs, ia[order_a], ib[order_b])


def test_mpc_coupled_meshes_equal_merged_mesh():
    """Two meshes tied at coincident interface nodes (u_b = u_a per dof) solve, through
    op_bc_mpc, to the merged mesh's displacements; a normal constraint on the x = 0 plane equals
    fixing u_x there."""
    E, NU = 210e3, 0.3
    (X, conn), (Xs, cs, ia, ib) = _two_bars()
    def solve(Xn, cn, mpcs, fixed):
        nd = Xn.shape[0] * 3
        ops = [(O.dofs(cn), O.element_stiffness(O.HEX8, Xn[cn], E, NU))]
        K = O.ops_assemble(ops, nd)
        T, smask = O.transformation(nd, mpcs, fixed) if mpcs else (sp.identity(nd, format="csr"), np.zeros(nd, bool))
        free = (~fixed & ~smask).astype(float)
        f = np.zeros(nd)
        tip = np.nonzero(Xn[:, 0] > Xn[:, 0].max() - 1e-9)[0]
        f[tip * 3 + 2] = -1.0
        A = O.bc_matrix((T.T @ K @ T).tocsr(), free)
        ub = spla.spsolve(A.tocsc(), free * (T.T @ f))
        return T @ ub
    import scipy.sparse as sp
    nd = X.shape[0] * 3
    fx = np.zeros(nd, bool)
    x0 = np.nonzero(X[:, 0] < 1e-9)[0]
    fx[(x0[:, None] * 3 + np.arange(3)).ravel()] = True
    u_ref = solve(X, conn, [], fx)
    nds = Xs.shape[0] * 3
    fs = np.zeros(nds, bool)
    x0s = np.nonzero(Xs[:, 0] < 1e-9)[0]
    fs[(x0s[:, None] * 3 + np.arange(3)).ravel()] = True
    tie = O.mpc_rows((ib[:, None] * 3 + np.arange(3)).ravel(), (ia[:, None] * 3 + np.arange(3)).ravel()[:, None],
                     np.ones((3 * len(ib), 1)))
    u_tied = solve(Xs, cs, [tie], fs)
    # compare node by node through the coordinates
    key = lambda P: np.lexsort((P[:, 2], P[:, 1], P[:, 0]))
    ur, ut = u_ref.reshape(-1, 3)[key(X)], u_tied.reshape(-1, 3)
    keep = np.ones(len(Xs), bool)
    keep[ib] = False
    ut = ut[keep][key(Xs[keep])]
    assert np.linalg.norm(ut - ur) <= 1e-10 * np.linalg.norm(ur)
    # normal constraint on x = 0 with n = (1, 0, 0): the same as fixing u_x there
    fx2 = np.zeros(nd, bool)
    fx2[x0 * 3 + 1] = fx2[x0 * 3 + 2] = True
    u_n = solve(X, conn, [O.normal_constraint(x0, (2.0, 0.0, 0.0))], fx2)
    fx3 = fx2.copy()
    fx3[x0 * 3] = True
    u_f = solve(X, conn, [], fx3)
    assert np.linalg.norm(u_n - u_f) <= 1e-10 * np.linalg.norm(u_f)


# ------------------------------------------------------------------ springs, contact, walls
def _fd_grad(fun, x, h=1e-6):
    x = np.array(x, float)
    g = np.zeros_like(x)
    for i in np.ndindex(x.shape):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (fun(xp) - fun(xm)) / (2 * h)
    return g


@pytest.mark.parametrize("kind", ["axial", "bushing", "ground"])
def test_spring_force_is_energy_gradient_and_stiffness_its_jacobian(kind):
    """spring_element (gfem/discrete.py:60-72,105-112): the force equals central differences of
    the energy (sc x), the stiffness central differences of the force; the axial tangent has the
    transverse part (1 - L0/L) k, which a small-rotation restatement would drop."""
    rng = np.random.default_rng(21)
    nn = 1 if kind == "ground" else 2
    Xe = rng.standard_normal((nn, 3)) * 2.0
    ue = rng.standard_normal((nn, 3)) * 0.3
    k = [3.0, 5.0, 7.0] if kind != "axial" else 4.0
    sc = 1.7
    fe, Ke = O.spring_element(kind, Xe, ue, k, sc)
    g = sc * _fd_grad(lambda u: O.spring_energy(kind, Xe, u, k), ue).ravel()
    assert np.allclose(fe, g, rtol=1e-7, atol=1e-9)
    J = np.stack([_fd_grad(lambda u, r=r: O.spring_element(kind, Xe, u, k, sc)[0][r], ue).ravel()
                  for r in range(3 * nn)])
    assert np.allclose(Ke, J, rtol=1e-6, atol=1e-8)
    assert np.allclose(Ke, Ke.T)
    if kind == "axial":
        e = Xe[1] + ue[1] - Xe[0] - ue[0]
        t = e / np.linalg.norm(e)
        assert abs(t @ Ke[3:, 3:] @ t - sc * 4.0) < 1e-12         # axial stiffness k along t
        assert not np.allclose(Ke[3:, 3:], sc * 4.0 * np.outer(t, t))


def test_contact_pair_gap_gradient_and_branches():
    """contact_pair (gfem/contact.py:19-36): dg == central differences of contact_gap; force
    -pn dg and tangent k dg dg^T when pressed, zero when open or inactive."""
    rng = np.random.default_rng(22)
    x = np.array([[0.2, 0.3, -0.05], [0, 0, 0], [1.0, 0.1, 0.02], [0.1, 0.9, -0.03]]) + 0.01 * rng.standard_normal((4, 3))
    k = 50.0
    f, K = O.contact_pair(x, k, 0.0, 1.0)
    dg = _fd_grad(O.contact_gap, x).ravel()
    g = O.contact_gap(x)
    assert g < 0
    assert np.allclose(f.ravel(), k * g * dg, rtol=1e-7, atol=1e-10)       # -pn dg with pn = -k g
    assert np.allclose(K, k * np.outer(dg, dg), rtol=1e-6, atol=1e-8)
    assert np.allclose(f.sum(0), 0, atol=1e-12)                            # internal pair force
    xo = x.copy()
    xo[0, 2] += 1.0                                                        # open: no force
    fo, Ko = O.contact_pair(xo, k, 0.0, 1.0)
    assert not fo.any() and not Ko.any()
    fi, Ki = O.contact_pair(x, k, 0.0, 0.0)                                # inactive
    assert not fi.any() and not Ki.any()
    xl = x.copy()
    xl[0, 2] += 0.05 - g                                                   # slightly open ...
    fl, _ = O.contact_pair(xl, k, 5.0, 1.0)                                # ... held by the multiplier
    gl = O.contact_gap(xl)
    assert gl > 0 and 5.0 - k * gl > 0
    assert np.allclose(fl.ravel(), -(5.0 - k * gl) * _fd_grad(O.contact_gap, xl).ravel(), rtol=1e-6)


def test_contact_select_and_wall_closed_forms():
    """contact_select == brute-force nearest clamped projection; wall forces push penetrating
    nodes back along the normal with k times the penetration."""
    rng = np.random.default_rng(23)
    X = rng.uniform(0, 1, (30, 3))
    X[:10, 2] = 0.0
    cand = np.stack([rng.choice(10, 3, replace=False) for _ in range(20 * 4)]).reshape(20, 4, 3)
    slave = np.arange(10, 30)
    u = 0.01 * rng.standard_normal(X.shape)
    nodes, active, g = O.contact_select(X, u, slave, cand, 0.5, 0.25)
    x = X + u

    def facet_dist(s, tri):
        # brute force: the nearest of dense barycentric samples of the facet
        A, B, C = x[tri]
        w = np.random.default_rng(s).dirichlet([1, 1, 1], 20000)
        P = w[:, :1] * A + w[:, 1:2] * B + w[:, 2:] * C
        return np.linalg.norm(P - x[slave[s]], axis=1).min()
    for s in range(20):
        dists = [facet_dist(s, cand[s, j]) for j in range(4)]
        assert facet_dist(s, nodes[s, 1:]) <= min(dists) + 5e-3, (s, dists)
        A, B, C = x[nodes[s, 1:]]
        n = np.cross(B - A, C - A)
        assert abs(g[s] - (x[slave[s]] - A) @ n / np.linalg.norm(n)) < 1e-12
        assert nodes[s, 0] == slave[s]
    fe, K = O.wall_evaluate(X, np.zeros_like(X) + [0, 0, -0.01], np.arange(5), [0, 0, 0], [0, 0, 2.0], 100.0)
    assert np.allclose(fe, [[0, 0, -1.0]] * 5)                 # -pn n with pn = k * 0.01
    assert np.allclose(K[0], np.diag([0, 0, 100.0]))


# ------------------------------------------------------------------ hyperelastic materials
def _hyper_params(kind):
    """Constants of each strain-energy material (small-strain mu ~ 77e3, bulk ~ 167e3 where the
    form allows) and its small-strain (mu, lambda)."""
    mu, lam = O.lame(E, NU)
    K = lam + 2 * mu / 3
    D = 2 / K
    p = dict(neohookean=(mu / 2, D), mooney_rivlin=(0.3 * mu, 0.2 * mu, D),
             yeoh=(mu / 2, -0.05 * mu, 0.01 * mu, D, 4 * D, 0.0),
             arruda_boyce=(mu, 2.5, D), gent=(mu, 8.0, D), svk=(mu, lam))[kind]
    return p


def _hyper_elements(et, seed=7):
    if et is O.TET10:
        X, c = _tet10()
    else:
        X, c = _brick(jitter=0.08, seed=seed)
    return X[c]


@pytest.mark.parametrize("kind", O.HYPER_KINDS)
def test_hyper_pk1_is_the_energy_gradient(kind):
    """Closed-form P == complex-step dW/dF (exact to rounding) at large deformations."""
    rng = np.random.default_rng(11)
    p = _hyper_params(kind)
    for _ in range(5):
        F = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0.3:
            continue
        P = O.hyper_pk1(kind, p, F)
        cs = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                Fc = F.astype(complex)
                Fc[i, j] += 1e-30j
                cs[i, j] = O.hyper_energy(kind, p, Fc).imag / 1e-30
        assert np.abs(P - cs).max() <= 1e-11 * np.abs(cs).max(), kind
    # a wrong invariant (I1 for I1bar) changes P
    if kind == "neohookean":
        F = np.diag([1.2, 0.9, 1.05])
        assert not np.allclose(O.hyper_pk1(kind, p, F), 2 * p[0] * F + 2 * (np.linalg.det(F) - 1) / p[1]
                               * np.linalg.det(F) * np.linalg.inv(F).T)


@pytest.mark.parametrize("kind", O.HYPER_KINDS)
@pytest.mark.parametrize("et", [O.HEX8, O.HEX8_PLAIN, O.TET10], ids=["hex8", "hex8_plain", "tet10"])
def test_hyper_force_is_the_energy_gradient(et, kind):
    """fe == complex-step gradient of the element strain energy: pins the F-bar pullback."""
    Xe = _hyper_elements(et)
    ue = 0.06 * np.random.default_rng(12).standard_normal(Xe.shape)
    p = _hyper_params(kind)
    fe = O.hyper_element_force(et, Xe, ue, kind, p)
    g = np.zeros_like(fe)
    for a in range(Xe.shape[1]):
        for k in range(3):
            uc = ue.astype(complex)
            uc[:, a, k] += 1e-30j
            g[:, a, k] = O.hyper_element_energy(et, Xe, uc, kind, p).imag / 1e-30
    assert np.abs(fe - g).max() <= 1e-10 * np.abs(g).max()


@pytest.mark.parametrize("kind", ["neohookean", "gent", "svk"])
@pytest.mark.parametrize("et", [O.HEX8, O.HEX8_PLAIN, O.TET10], ids=["hex8", "hex8_plain", "tet10"])
def test_hyper_stiffness_pins(et, kind):
    """Ke symmetric, == central differences of the force; at u = 0 == the linear element
    stiffness (B-bar for C3D8) with the material's small-strain constants."""
    Xe = _hyper_elements(et)
    p = _hyper_params(kind)
    ue = 0.05 * np.random.default_rng(13).standard_normal(Xe.shape)
    K = O.hyper_element_stiffness(et, Xe, ue, kind, p)
    assert np.abs(K - np.swapaxes(K, 1, 2)).max() <= 1e-9 * np.abs(K).max()
    h = 1e-6
    nn = Xe.shape[1]
    for col in (0, 4, 3 * nn - 1):
        up, um = ue.copy(), ue.copy()
        up.reshape(len(ue), -1)[:, col] += h
        um.reshape(len(ue), -1)[:, col] -= h
        fd = (O.hyper_element_force(et, Xe, up, kind, p) - O.hyper_element_force(et, Xe, um, kind, p)) / (2 * h)
        assert np.abs(K[:, :, col] - fd.reshape(len(ue), -1)).max() <= 1e-6 * np.abs(K).max()
    K0 = O.hyper_element_stiffness(et, Xe, np.zeros_like(ue), kind, p)
    assert np.abs(K0 - O.element_stiffness(et, Xe, E, NU)).max() <= 1e-10 * np.abs(K0).max()


# ------------------------------------------------------------------ buckling
def test_element_cauchy_closed_forms():
    """Uniaxial homogeneous strain (lateral contraction -nu eps): sigma = E eps e_x e_x; a
    homogeneous stretch of a NeoHookean block: 2 C10 J^-5/3 dev(B) + 2 (J - 1)/D1 I."""
    X, c = _brick(jitter=0.05, seed=3)
    Xe = X[c]
    eps = 1e-3
    ue = np.stack([eps * Xe[..., 0], -NU * eps * Xe[..., 1], -NU * eps * Xe[..., 2]], -1)
    for et in (O.HEX8, O.HEX8_PLAIN):
        s = O.element_cauchy(et, Xe, ue, ("linear", (E, NU)))[0]
        want = np.zeros((3, 3))
        want[0, 0] = E * eps
        assert np.abs(s - want).max() <= 1e-9 * E * eps
    F = np.array([[1.3, 0.1, 0.0], [0.05, 0.9, 0.02], [0.0, -0.04, 1.1]])
    ue = np.einsum("ij,eaj->eai", F - np.eye(3), Xe)
    p = _hyper_params("neohookean")
    J = np.linalg.det(F)
    B = F @ F.T
    want = 2 * p[0] * J ** (-5 / 3) * (B - np.trace(B) / 3 * np.eye(3)) + 2 * (J - 1) / p[1] * np.eye(3)
    for et in (O.HEX8, O.HEX8_PLAIN):
        s = O.element_cauchy(et, Xe, ue, ("neohookean", p))[0]
        assert np.abs(s - want).max() <= 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("et", [O.HEX8, O.TET10], ids=["hex8", "tet10"])
def test_geo_stiff_is_the_initial_stress_work(et):
    """u^T KG u == sum_g vol tr(H sig H^T), H = grad u (the second-order work of the initial
    stress) for a non-symmetric test stress; rigid translations give zero."""
    Xe = _hyper_elements(et)
    rng = np.random.default_rng(31)
    sig = rng.standard_normal((len(Xe), 3, 3))
    K = O.geo_stiff(et, Xe, sig)
    dNdX, vol = O.geom(et, Xe)
    for _ in range(3):
        ue = rng.standard_normal(Xe.shape)
        H = np.einsum("eai,egaj->egij", ue, dNdX)
        w = np.einsum("eg,egij,ejk,egik->e", vol, H, sig, H)
        q = np.einsum("ei,eij,ej->e", ue.reshape(len(Xe), -1), K, ue.reshape(len(Xe), -1))
        assert np.abs(q - w).max() <= 1e-12 * np.abs(w).max()
    t = np.zeros(Xe.shape)
    t[..., 1] = 1.0
    assert np.abs(np.einsum("eij,ej->ei", K, t.reshape(len(Xe), -1))).max() <= 1e-12 * np.abs(K).max()


def test_buckling_factor_of_a_clamped_column_is_euler():
    """A clamped column (20 x 1 x 1) under end compression: the first load factors of two meshes,
    Richardson-extrapolated in h^2, within 1% of Euler's pi^2 E I / (4 L^2); a pair of equal
    factors (two bending planes)."""
    L = 20.0
    lam1 = []
    for n in (3, 4):
        X, c = O.box_hex(L, 1.0, 1.0, 10 * n, n, n)
        nd = 3 * len(X)
        K0 = O.ops_assemble([(O.dofs(c), O.element_stiffness(O.HEX8, X[c], E, NU))], nd)
        fix = np.nonzero(X[:, 0] < 1e-9)[0]
        tip = np.nonzero(X[:, 0] > L - 1e-9)[0]
        free = np.ones(nd)
        free[(fix[:, None] * 3 + np.arange(3)).ravel()] = 0
        f = np.zeros(nd)
        f[tip * 3] = -1.0 / len(tip)
        u = spla.spsolve(O.bc_matrix(K0, free).tocsc(), free * f)
        sig = O.element_cauchy(O.HEX8, X[c], u.reshape(-1, 3)[c], ("linear", (E, NU)))
        KG = O.ops_assemble([(O.dofs(c), O.geo_stiff(O.HEX8, X[c], sig))], nd)
        lam, _ = O.buckling_factors(K0, KG, free, 3)
        assert abs(lam[1] / lam[0] - 1) < 1e-6 and lam[2] > 5 * lam[0]
        lam1.append(lam[0])
    ext = (16 * lam1[1] - 9 * lam1[0]) / 7
    euler = np.pi ** 2 * E / 12 / (4 * L ** 2)
    assert abs(ext / euler - 1) < 0.01, (lam1, ext / euler)


# ------------------------------------------------------------------ thermal and eigenstrain fields
@pytest.mark.parametrize("et", [O.HEX8, O.HEX8_PLAIN], ids=["hex8", "hex8_plain"])
def test_field_strain_closed_forms(et):
    """Uniform dT with u = alpha dT X (free expansion): zero element forces; u = 0 (clamped): the
    x = L face carries -3 K alpha dT A in x; an eigenstrain eps0 I == the thermal strain
    alpha dT = eps0; a sign error or a missing interpolation fails one of them."""
    X, c = O.box_hex(2.0, 1.0, 1.0, 4, 2, 2)
    X = X.copy()
    X[:, 0] += 0.2 * X[:, 2]                        # sheared: parallelepipeds (B-bar exact)
    Xe = X[c]
    alpha, dT = 1.2e-5, 80.0
    Te = np.full(c.shape, 20.0 + dT)
    T0e = np.full(c.shape, 20.0)
    fe = O.element_force(et, Xe, alpha * dT * Xe, E, NU, Te=Te, T0e=T0e, alpha=alpha)
    assert np.abs(fe).max() <= 1e-12 * E * alpha * dT
    fe = O.element_force(et, Xe, np.zeros_like(Xe), E, NU, Te=Te, T0e=T0e, alpha=alpha)
    f = np.zeros((len(X), 3))
    np.add.at(f, c, fe)
    mu, lam = O.lame(E, NU)
    face = np.nonzero(np.isclose(X[:, 0] - 0.2 * X[:, 2], 2.0))[0]
    assert abs(f[face, 0].sum() - (-(3 * lam + 2 * mu) * alpha * dT * 1.0 * 1.0)) <= 1e-9 * E * alpha * dT
    eps0 = alpha * dT
    eig = np.tile(eps0 * np.eye(3).ravel(), (len(c), 1))
    fe_e = O.element_force(et, Xe, np.zeros_like(Xe), E, NU, eig=eig)
    assert np.abs(fe_e - fe).max() <= 1e-12 * np.abs(fe).max()


# ----------------------------------------------------------------------------- SA set-up
def _chain(n):
    import scipy.sparse as sp
    return sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(n, n)).tocsr()


def test_sa_standard_aggregation_published_and_hand_cases():
    """PyAMG's own examples for standard_aggregation (its docstring: the 4-vertex 1D Poisson
    matrix gives two aggregates; a first vertex with no neighbour stays out), plus hand-worked
    cases: a 7-node chain (seeds 0, 3, 6 in pass 1), a 3x3 grid (two corner seeds, the two
    left-over nodes join in pass 2) and directed graphs that reach pass 3, including the
    published loop's renumbering of a neighbour it marked (aggregate id 0 reads as unmarked)."""
    import scipy.sparse as sp
    agg, k = O.sa_standard_aggregation(O.sa_strength(_chain(4), 1))
    assert k == 2 and agg.tolist() == [0, 0, 1, 1]
    A = sp.csr_matrix(np.array([[1.0, 0, 0], [0, 1, 1], [0, 1, 1]]))
    agg, k = O.sa_standard_aggregation(O.sa_strength(A, 1))
    assert k == 1 and agg.tolist() == [-1, 0, 0]
    agg, k = O.sa_standard_aggregation(O.sa_strength(_chain(7), 1))
    assert k == 3 and agg.tolist() == [0, 0, 1, 1, 1, 2, 2]
    grid = [[1, 3], [0, 2, 4], [1, 5], [0, 4, 6], [1, 3, 5, 7], [2, 4, 8], [3, 7], [4, 6, 8], [5, 7]]
    agg, k = O.sa_standard_aggregation(grid)
    assert k == 2 and agg.tolist() == [0, 0, 1, 0, 1, 1, 0, 1, 1]
    agg, k = O.sa_standard_aggregation([[], [0]])
    assert k == 1 and agg.tolist() == [-1, 0]
    agg, k = O.sa_standard_aggregation([[], [0, 2], [0], [0]])
    assert k == 3 and agg.tolist() == [-1, 0, 1, 2]
    # a zero block is not a connection (clamped nodes of bc_matrix are isolated)
    A = _chain(4).tolil()
    A[0, 1] = A[1, 0] = 0.0
    assert O.sa_strength(A.tocsr(), 1)[0] == []


def test_sa_fit_candidates_published_and_rank_deficient():
    """PyAMG's fit_candidates examples (its docstring: two aggregates of two nodes, the
    constant candidate -> Q = 1/sqrt(2), R = sqrt(2); constant + linear -> the listed Q and R),
    and a one-node aggregate with the six rigid-body candidates (rank 3): three orthonormal
    columns, three zero columns, T R == B."""
    agg = np.array([0, 0, 1, 1])
    T, R = O.sa_fit_candidates(agg, 2, np.ones((4, 1)), 1)
    s = np.sqrt(0.5)
    assert np.allclose(T.toarray(), [[s, 0], [s, 0], [0, s], [0, s]], atol=1e-15)
    assert np.allclose(R, [[np.sqrt(2)], [np.sqrt(2)]], atol=1e-15)
    T, R = O.sa_fit_candidates(agg, 2, np.array([[1.0, 0], [1, 1], [1, 2], [1, 3]]), 1)
    assert np.allclose(T.toarray(), [[s, -s, 0, 0], [s, s, 0, 0], [0, 0, s, -s], [0, 0, s, s]], atol=1e-15)
    assert np.allclose(R, [[np.sqrt(2), s], [0, s], [np.sqrt(2), 2.5 * np.sqrt(2)], [0, s]], atol=1e-14)
    B = O.rigid_body_modes(np.array([[0.0, 0, 0], [1.0, 0.5, 0.2]]))[:3]
    T, R = O.sa_fit_candidates(np.array([0]), 1, B, 3)
    Q = T.toarray()
    nz = np.abs(Q).sum(0) > 0
    assert nz.sum() == 3 and np.allclose(Q[:, nz].T @ Q[:, nz], np.eye(3), atol=1e-14)
    assert np.allclose(Q @ R, B, atol=1e-14)


def test_sa_smooth_prolongator_closed_form():
    """P = T - (4/3) D^-1 A T with Gershgorin row sums D = (3, 4, 4, 3) on the 4-node chain,
    aggregates {0,1} and {2,3}, constant candidate: by hand P = (1/sqrt 2) [[5/9, 0],
    [2/3, 1/3], [1/3, 2/3], [0, 5/9]] (the diagonal D = 2 would give 1/3 in the first row)."""
    T, _ = O.sa_fit_candidates(np.array([0, 0, 1, 1]), 2, np.ones((4, 1)), 1)
    P = O.sa_smooth_prolongator(_chain(4), T).toarray()
    a = np.sqrt(0.5)
    assert np.allclose(P, a * np.array([[5 / 9, 0], [2 / 3, 1 / 3], [1 / 3, 2 / 3], [0, 5 / 9]]), atol=1e-15)


def test_sa_hierarchy_coarse_operators_annihilate_the_coarse_candidates():
    """On a free-floating elastic box A B = 0 for the rigid-body modes B, and T Bc = B, so
    P Bc = B - omega D^-1 A B = B and A_c Bc = P^T A B = 0 at every level (a dropped smoothing
    term, a wrong Bc or a non-Galerkin coarse operator breaks it). Also: aggregates cover every
    node, the level loop stops at max_coarse block rows."""
    X, c = O.box_hex(4.0, 2.0, 2.0, 6, 3, 3)
    X = X + np.random.default_rng(3).uniform(-0.05, 0.05, X.shape)
    A = O.ops_assemble([(O.dofs(c), O.element_stiffness(O.HEX8, X[c], E, NU))], 3 * len(X))
    B = O.rigid_body_modes(X)
    levels, Ac = O.sa_hierarchy(A, B, 3, max_coarse=4)
    assert len(levels) >= 2 and Ac.shape[0] // 6 <= 4
    Bl = B
    for lv in levels:
        assert (lv["agg"] >= 0).all()
        P, Bc = lv["P"], lv["Bc"]
        assert np.abs(P @ Bc - Bl).max() <= 1e-9 * np.abs(Bl).max()
        Al = lv["A"]
        Acl = P.T @ Al @ P
        assert np.abs(Acl @ Bc).max() <= 1e-9 * abs(Al).max() * np.abs(Bc).max()
        Bl = Bc
    prev = [lv["A"].shape[0] // bs for lv, bs in zip(levels, [3] + [6] * 10)]
    assert all(nb > 4 for nb in prev)


def test_sa_coarse_operator_is_galerkin_for_its_prolongator():
    """On the clamped (SPD) system the two-level correction E = I - P A_c^-1 P^T A is an
    A-orthogonal projector (E E = E, A E symmetric) exactly when A_c = P^T A P for the P that
    prolongs (an A_c formed from the tentative T, or with R != P^T, is not)."""
    X, c = O.box_hex(4.0, 2.0, 2.0, 6, 3, 3)
    A = O.ops_assemble([(O.dofs(c), O.element_stiffness(O.HEX8, X[c], E, NU))], 3 * len(X))
    free = np.ones(3 * len(X))
    free[np.repeat(X[:, 0] < 1e-9, 3)] = 0.0
    Abc = O.bc_matrix(A, free)
    levels, _ = O.sa_hierarchy(Abc, O.rigid_body_modes(X), 3, max_coarse=4)
    P = levels[0]["P"].toarray()
    Ad = Abc.toarray()
    Ac = levels[1]["A"].toarray() if len(levels) > 1 else None
    Ac = P.T @ Ad @ P if Ac is None else Ac
    Ec = np.eye(len(Ad)) - P @ np.linalg.solve(Ac, P.T @ Ad)
    assert np.abs(Ec @ Ec - Ec).max() <= 1e-9
    assert np.abs(Ad @ Ec - (Ad @ Ec).T).max() <= 1e-9 * np.abs(Ad).max()
    T = levels[0]["T"].toarray()
    Et = np.eye(len(Ad)) - P @ np.linalg.solve(T.T @ Ad @ T, P.T @ Ad)
    assert np.abs(Et @ Et - Et).max() > 1e-3
