"""Block PCG at the batched workload's own shape: 64 load cases (multi_load_cases,
gfem/optimize.py:116-143) of seeded Gaussian top-face forces on a jittered C3D8 box, solved with
every variant of the batched product pinned in turn (exfem_op_set_multi_shape: EbE MFMA slot
shapes, patch node caps, the assembled operator), for Jacobi FP64 and FP32 and the batched
SA-AMG V-cycle (FP64 cycle, FP32 cycle in the FP64 solve). Every case of every run is compared
with a sparse direct solve of the oracle's system: FP64 solves within 1e-8; the all-FP32 solve
within 2e-4, the accuracy of its FP32 operator (cond(K) x eps32 on this mesh), stated per run."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu
NCASE = 64
NVARIANTS = 11


@pytest.fixture(scope="module")
def problem():
    from exfem_amd import workloads
    p = workloads.cantilever(12, 4, 4, jitter=0.1, seed=31)
    F = workloads.top_face_loads(p, NCASE)
    ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], p.E, p.nu))]
    free = p.free.astype(float)
    lu = spla.splu(O.bc_matrix(O.ops_assemble(ops, p.ndof), free).tocsc())
    U0 = np.stack([lu.solve(free * F[k]) for k in range(NCASE)])
    return p, F, U0


def _max_rel(U, U0):
    U = U.detach().cpu().numpy().astype(float)
    return float((np.linalg.norm(U - U0, axis=1) / np.linalg.norm(U0, axis=1)).max())


@pytest.mark.parametrize("precond,dtype", [("jacobi", 0), ("jacobi", 1), ("amg", 0), ("amg32", 0)])
def test_64_case_block_pcg_every_variant_matches_direct(gpu, problem, precond, dtype):
    import exfem_amd as X
    from exfem_amd import _lib as L
    from exfem_amd.linalg import AMG
    from exfem_amd.optimize import multi_load_cases
    p, F, U0 = problem
    m = p.model()
    fix = np.nonzero(p.X[:, 0] < 1e-9)[0]
    step = X.Step(bcs=[(fix, d, 0.0) for d in range(3)])
    op = m.operator()
    amg = None
    if precond != "jacobi":
        op.set_ke(0, m.element_stiffness()[0])
        op.set_bc(p.free)
        amg = AMG(op.assemble(bc=True), m.X, max_coarse=100)
    cycle_dt = L.F32 if (precond == "amg32" or dtype == L.F32) else L.F64
    tol = 1e-8 if dtype == L.F64 else 1e-6
    bound = 1e-8 if dtype == L.F64 else 2e-4
    ran, errs = 0, {}
    product_dts = sorted({cycle_dt, dtype})     # the batched products this solve runs
    for i in range(NVARIANTS):
        pinned = []
        for dt in product_dts:
            try:
                op.set_multi_shape(NCASE, i, dtype=dt)
                pinned.append(dt)
            except L.ExfemError:
                pass                      # a patch variant beyond the LDS at this dtype
        if not pinned:
            continue
        U, info = multi_load_cases(m, step, F, cg_tol=tol, dtype=dtype, return_info=True,
                                   precond=precond, amg=amg)
        for dt in pinned:
            assert op.multi_shape(NCASE, dtype=dt)["index"] == i
        assert info["status"] == 0 and (info["relres"] <= tol).all(), (i, info["relres"].max())
        errs[i] = _max_rel(U, U0)
        assert errs[i] <= bound, (precond, dtype, i, errs[i])
        if precond != "jacobi":
            assert info["iters"].max() <= 40, info["iters"].max()
        ran += 1
    for dt in product_dts:
        op.set_multi_shape(NCASE, None, dtype=dt)
    assert ran >= 5, errs
    print(f"{precond} dtype {dtype}: max rel error vs FP64 direct per variant {errs}")


@pytest.mark.parametrize("kind", ["hex", "tet10"])
def test_patch_staged_product_equals_the_four_case_form(gpu, kind):
    """The batched product over node patches with LDS-staged x rows (variant 10) is bitwise the
    four-cases-per-lane product on the assembled operator (variant 9), FP64 and FP32, with and
    without the boundary mask, for 16, 32 and 64 cases, and within 1e-14 (FP64) of the
    single-vector products."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    if kind == "hex":
        p = workloads.cantilever(14, 6, 5, jitter=0.1, seed=8)
        f = None
    else:
        p = workloads.tet10_box(3, 3, 6, seed=4)
        f = dict(scale=[p.scale])
    m = p.model()
    op = m.operator(f)
    for dt in (L.F64, L.F32):
        op.set_ke(0, m.element_stiffness(storage=1, dtype=dt, fields=f)[0])
        op.set_bc(p.free)
        for R in (16, 32, 64):
            U = torch.as_tensor(np.random.default_rng(R).standard_normal((R, p.ndof)), device="cuda")
            U = U if dt == L.F64 else U.float()
            for bc in (False, True):
                op.set_multi_shape(R, 9, dtype=dt)
                Y9 = op.matvec_multi(U, bc=bc, dtype=dt)
                op.set_multi_shape(R, 10, dtype=dt)
                Y10 = op.matvec_multi(U, bc=bc, dtype=dt)
                assert torch.equal(Y9, Y10), (dt, R, bc)
                if dt == L.F64:
                    for r in (0, R - 1):
                        y = op.matvec(U[r], bc=bc)
                        assert float(torch.linalg.norm(Y10[r] - y)) <= 1e-14 * float(torch.linalg.norm(y))
        op.set_multi_shape(64, None, dtype=dt)
