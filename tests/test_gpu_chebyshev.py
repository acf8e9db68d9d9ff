"""Device Chebyshev paths against the oracle, one application at a time (through the C ABI):
the Chebyshev-Jacobi preconditioner of _pcg_fn (gfem/solvers.py:289-307), the level smoother of
the V-cycle (gfem/linalg.py:161-175), the _lmax power estimate (gfem/solvers.py:328-342) and the
device sparse product of csr_mv (gfem/linalg.py:51-77).

The oracle functions these compare with are pinned by closed forms in tests/test_oracle.py
(Chebyshev polynomials T_3 / T_7 on known spectra, power-iteration ratios from eigenvalue
powers). Tolerances: FP64 1e-10 relative (norm-wise), FP32 1e-5.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
E, NU = 210000.0, 0.3


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _np(t):
    return t.detach().cpu().numpy()


def _problem(nx=10, ny=3, nz=3, seed=5):
    """Jittered C3D8 cantilever with its device operator (sym FP64 and FP32 element matrices,
    geometry for the matrix-free form, boundary mask) and the oracle's masked operator."""
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    p = workloads.cantilever(nx, ny, nz, jitter=0.1, seed=seed)
    m = p.model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness(storage=L.KE_SYM)[0])
    op.set_ke(0, m.element_stiffness(storage=L.KE_SYM, dtype=L.F32)[0])
    op.set_bc(p.free)
    ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], E, NU))]
    free = p.free.astype(float)
    A0 = O.op_bc(lambda w: O.ops_matvec(ops, w, p.ndof), free)
    dinv0 = O.jacobi_inv(O.ops_diagonal(ops, p.ndof), free)
    return p, m, op, ops, free, A0, dinv0


@pytest.mark.parametrize("matrix_free", [False, True])
def test_lmax_power_matches_oracle(gpu, matrix_free):
    import torch
    from exfem_amd.linalg import BcOperator, lmax_power
    p, m, op, ops, free, A0, dinv0 = _problem()
    lm = lmax_power(BcOperator(op, matrix_free=matrix_free), torch.as_tensor(dinv0, device="cuda"))
    lm0 = O.lmax_power(A0, dinv0, free)
    assert abs(lm - lm0) <= 1e-11 * lm0, (lm, lm0)
    # the reference's iteration count matters at this spectrum: 24 steps give another value
    lm24 = lmax_power(BcOperator(op, matrix_free=matrix_free), torch.as_tensor(dinv0, device="cuda"), 24)
    assert abs(lm24 - lm0) > 1e-9 * lm0


@pytest.mark.parametrize("matrix_free", [False, True])
def test_chebyshev_preconditioner_matches_oracle(gpu, matrix_free):
    """One EXFEM_PC_CHEBY application (exfem_pcg_precond) == O.cheby_precond on the same masked
    operator, dinv and lmax, FP64 (1e-10) and on FP32 vectors (1e-5); Jacobi and none too."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import BcOperator, precond_apply
    p, m, op, ops, free, A0, dinv0 = _problem(seed=6)
    lmax = O.lmax_power(A0, dinv0, free)
    dd = torch.as_tensor(dinv0, device="cuda")
    rng = np.random.default_rng(7)
    for trial in range(3):
        r = rng.standard_normal(p.ndof)
        z0 = O.cheby_precond(A0, dinv0, lmax)(r)
        z = precond_apply(L.PC_CHEBY, r, op=op, pinv=dd, lmax=lmax, matrix_free=matrix_free)
        assert rel(_np(z), z0) <= 1e-10, (trial, rel(_np(z), z0))
    z32 = precond_apply(L.PC_CHEBY, torch.as_tensor(r, device="cuda").float(), op=op, pinv=dd,
                        lmax=lmax, matrix_free=matrix_free, dtype=L.F32)
    assert z32.dtype == torch.float32 and rel(_np(z32).astype(float), z0) <= 1e-5
    assert rel(_np(precond_apply(L.PC_JACOBI, r, pinv=dd)), dinv0 * r) <= 1e-15
    assert np.array_equal(_np(precond_apply(L.PC_NONE, r)), r)
    # a different interval gives a different preconditioner, far outside the tolerance
    z_wrong = O.cheby_precond(A0, dinv0, lmax * 3.0)(r)
    assert rel(_np(z), z_wrong) > 1e-4


def test_chebyshev_solver_backend_uses_device_lmax(gpu):
    """StaticSolver(linear_solver='cheby') with the device _lmax and Chebyshev application
    solves to the FP64 direct solution."""
    import scipy.sparse.linalg as spla
    import exfem_amd as X
    p, m, op, ops, free, A0, dinv0 = _problem(seed=8)
    fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
    st = X.Step(bcs=[(fix, d, 0.0) for d in range(3)], dloads=[(p.f,)])
    u, _, info = X.StaticSolver(m, verbose=False, linear_solver="cheby", cg_tol=1e-10).solve(st)
    ud = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, p.ndof), free).tocsc(), free * p.f)
    assert rel(_np(u), ud) <= 1e-8


@pytest.mark.parametrize("matrix_free", [False, True])
def test_level0_smoother_matches_oracle(gpu, matrix_free):
    """One smoother call on level 0 (exfem_amg_smooth, the EbE operator with the boundary mask)
    == O.cheb with the same dinv0 and rho: from x = 0 (the V-cycle's pre-smoother, which skips
    A*0) and from a given x; FP64 1e-10 and FP32 1e-5."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import AMG, BcOperator
    p, m, op, ops, free, A0, dinv0 = _problem(seed=9)
    amg = AMG(O.bc_matrix(O.ops_assemble(ops, p.ndof), free), p.X, max_coarse=60)
    rho = amg.levels_host[0]["rho"]
    dd = torch.as_tensor(dinv0, device="cuda")
    Aop = BcOperator(op, matrix_free=matrix_free)
    rng = np.random.default_rng(10)
    b = free * rng.standard_normal(p.ndof)
    x = rng.standard_normal(p.ndof)
    x_from0 = O.cheb(A0, dinv0, rho, np.zeros(p.ndof), b, deg=2)
    x_fromx = O.cheb(A0, dinv0, rho, x, b, deg=2)
    assert rel(_np(amg.smooth(0, b, None, A0=Aop, dinv0=dd)), x_from0) <= 1e-10
    assert rel(_np(amg.smooth(0, b, x, A0=Aop, dinv0=dd)), x_fromx) <= 1e-10
    amg.set_precision(L.F32)
    Aop32 = BcOperator(op, matrix_free=matrix_free, dtype=L.F32)
    x32 = amg.smooth(0, b, x, A0=Aop32, dinv0=dd, dtype=L.F32)
    assert x32.dtype == torch.float32 and rel(_np(x32).astype(float), x_fromx) <= 1e-5
    # degree 2 (the reference) is not degree 3
    lo10 = O.cheb(A0, dinv0, rho, x, b, deg=3)
    assert rel(_np(amg.smooth(0, b, x, A0=Aop, dinv0=dd)), lo10) > 1e-4


def test_coarse_level_smoother_matches_oracle(gpu):
    """Smoothing on level 1 (its stored matrix and dinv, ELL or CSR as uploaded) and on level 0
    from the stored level-0 matrix == O.cheb, FP64 1e-10, FP32 1e-5."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import AMG
    p, m, op, ops, free, A0, dinv0 = _problem(24, 6, 6, seed=11)
    amg = AMG(O.bc_matrix(O.ops_assemble(ops, p.ndof), free), p.X, max_coarse=40)
    assert amg.nlevels >= 3
    rng = np.random.default_rng(12)
    for lvl in (0, 1):
        lv = amg.levels_host[lvl]
        A = sp.csr_matrix(lv["A"])
        b = rng.standard_normal(A.shape[0])
        x = rng.standard_normal(A.shape[0])
        want = O.cheb(lambda v: A @ v, lv["dinv"], lv["rho"], x, b, deg=2)
        assert rel(_np(amg.smooth(lvl, b, x)), want) <= 1e-10, lvl
        want0 = O.cheb(lambda v: A @ v, lv["dinv"], lv["rho"], np.zeros_like(b), b, deg=2)
        assert rel(_np(amg.smooth(lvl, b)), want0) <= 1e-10, lvl
    amg.set_precision(L.F32)
    got = amg.smooth(1, b, x, dtype=L.F32)
    assert got.dtype == torch.float32 and rel(_np(got).astype(float), want) <= 1e-5


@pytest.mark.parametrize("form", ["ell", "csr"])
def test_csr_mv_matches_scipy(gpu, form):
    """csr_to_device / csr_mv through the library: ELL (rows of similar length) and CSR (one long
    row), FP64 1e-14 and bitwise reproducible, FP32 1e-6."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import csr_mv, csr_to_device
    rng = np.random.default_rng(13)
    n = 3000
    A = sp.diags([rng.standard_normal(n - abs(k)) for k in (-40, -3, -1, 0, 1, 3, 40)],
                 [-40, -3, -1, 0, 1, 3, 40], format="csr")
    if form == "csr":
        A = sp.lil_matrix(A)
        A[5, :] = rng.standard_normal(n)
        A = sp.csr_matrix(A)
    x = rng.standard_normal(n)
    M = csr_to_device(A)
    assert ("ell_data" in M) == (form == "ell")
    y = csr_mv(M, x, n)
    assert rel(_np(y), A @ x) <= 1e-14
    assert torch.equal(y, csr_mv(M, x, n))
    M32 = csr_to_device(A, dtype=L.F32)
    assert rel(_np(csr_mv(M32, x, n)).astype(float), A @ x) <= 1e-6


@pytest.mark.parametrize("matrix_free", [False, True])
def test_fp32_chebyshev_preconditioner_in_fp64_pcg(gpu, matrix_free):
    """EXFEM_PC_CHEBY_MIXED: one application (FP32 polynomial, FP64 vectors) within 1e-5 of
    O.cheby_precond; inside FP64 PCG it takes at most 25% more iterations than the FP64 polynomial
    (the FP32 rounding perturbs the preconditioner by about 1e-7, which costs CG a little
    conjugacy at tol 1e-11) and solves to the direct solution (1e-8), stored FP32 matrices or FP32
    matrix-free."""
    import torch
    import scipy.sparse.linalg as spla
    from exfem_amd import _lib as L
    from exfem_amd.linalg import BcOperator, Chebyshev, lmax_power, pcg_device, precond_apply
    p, m, op, ops, free, A0, dinv0 = _problem(16, 4, 4, seed=14)
    dd = torch.as_tensor(dinv0, device="cuda")
    A = BcOperator(op, matrix_free=matrix_free)
    lmax = lmax_power(A, dd)
    r = np.random.default_rng(15).standard_normal(p.ndof)
    z0 = O.cheby_precond(A0, dinv0, lmax)(r)
    z = precond_apply(L.PC_CHEBY_MIXED, r, op=op, pinv=dd, lmax=lmax, matrix_free=matrix_free)
    assert z.dtype == torch.float64 and rel(_np(z), z0) <= 1e-5
    b = free * p.f
    x64, k64, _, s64 = pcg_device(A, b, Chebyshev(dd, lmax, A), tol=1e-11, maxiter=5000)
    x32, k32, _, s32 = pcg_device(A, b, Chebyshev(dd, lmax, A, fp32=True), tol=1e-11, maxiter=5000)
    ud = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, p.ndof), free).tocsc(), b)
    assert s64 == s32 == 0 and k32 <= 1.25 * k64 + 2, (k64, k32)
    assert rel(_np(x32), ud) <= 1e-8 and rel(_np(x64), ud) <= 1e-8
