"""The SA set-up with its products on the device (AMG.from_operator, exfem_sa_dev_*) against the
oracle's restatement of PyAMG's set-up (gfem/linalg.py:103-136) and the host builder.

Tolerances: aggregates identical; P, R and the level matrices <= 1e-12 relative (max norm);
rho <= 1e-12 relative; V-cycles <= 1e-10; solutions <= 1e-8 vs the FP64 direct solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import gfem_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def mrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _np(t):
    return t.detach().cpu().numpy()


def _problem(kind):
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    if kind in ("hex", "hex2"):
        p = workloads.cantilever(*((14, 5, 5) if kind == "hex" else (30, 8, 8)), jitter=0.1, seed=4)
        ops = [(O.dofs(p.conn), O.element_stiffness(O.HEX8, p.X[p.conn], p.E, p.nu))]
    else:
        p = workloads.tet10_box(3, 3, 7, seed=5)
        ops = [(O.dofs(p.conn), O.element_stiffness(O.TET10, p.X[p.conn], p.E, p.nu, scale=p.scale))]
    m = p.model()
    f = dict(scale=[p.scale]) if p.scale is not None else None
    op = m.operator(f)
    op.set_ke(0, m.element_stiffness(fields=f)[0])
    op.set_ke(0, m.element_stiffness(fields=f, dtype=L.F32)[0])
    op.set_bc(p.free)
    free = p.free.astype(float)
    A = O.bc_matrix(O.ops_assemble(ops, p.ndof), free)
    return p, op, A, ops, free


@pytest.mark.parametrize("kind", ["hex", "tet10"])
def test_device_sa_levels_equal_the_oracle_and_the_host_builder(gpu, kind):
    from exfem_amd.linalg import AMG
    p, op, A, ops, free = _problem(kind)
    assert mrel(op.assemble(bc=True).toarray(), A.toarray()) <= 1e-13
    amg = AMG.from_operator(op, p.X, max_coarse=8, keep_levels=True)
    ref, Ac = O.sa_hierarchy(A, O.rigid_body_modes(p.X), 3, max_coarse=8)
    host = AMG(A, p.X, max_coarse=8)
    assert len(amg.levels_host) == len(ref) == len(host.levels_host) >= 2
    assert amg.sizes == host.sizes
    for lv, (d, r, h) in enumerate(zip(amg.levels_host, ref, host.levels_host)):
        assert np.array_equal(d["agg"], r["agg"]) and np.array_equal(d["agg"], h["agg"]), lv
        assert mrel(d["P"].toarray(), r["P"].toarray()) <= 1e-12, lv
        assert mrel(d["R"].toarray(), r["P"].T.toarray()) <= 1e-12, lv
        assert mrel(sp.csr_matrix(d["A"]).toarray(), r["A"].toarray()) <= 1e-12, lv
        assert abs(d["rho"] - h["rho"]) <= 1e-12 * h["rho"], lv
        assert abs(d["rho"] - O.rho_estimate(r["A"], d["dinv"])) <= 1e-12 * d["rho"], lv
        assert mrel(d["dinv"], h["dinv"]) <= 1e-13
    Lr = O.coarse_factor(Ac.toarray())
    assert mrel(amg.Lc @ amg.Lc.T, Lr @ Lr.T) <= 1e-12
    # operator complexity counts stored blocks (the masked zero blocks of clamped nodes included)
    assert abs(amg.complexity - host.complexity) <= 0.02 * host.complexity


@pytest.mark.parametrize("kind", ["hex", "tet10"])
def test_device_sa_vcycle_and_solves(gpu, kind):
    """V-cycles of the device-built hierarchy == the oracle's V-cycle on the oracle's hierarchy
    (level 0 on the EbE operator and on the stored matrix, single and batched); AMG-PCG with it
    == the direct solve, FP64 and with the FP32 cycle."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd.linalg import AMG, AMGPrecond, BcOperator, pcg_device
    p, op, A, ops, free = _problem(kind)
    amg = AMG.from_operator(op, p.X, max_coarse=8)
    ref, Ac = O.sa_hierarchy(A, O.rigid_body_modes(p.X), 3, max_coarse=8)
    lv = []
    for r in ref:
        d = r["A"].diagonal()
        dinv = np.where(np.abs(d) > 0, 1.0 / np.where(d == 0, 1, d), 0.0)
        lv.append(dict(A=r["A"], dinv=dinv, rho=O.rho_estimate(r["A"], dinv), P=r["P"], R=r["P"].T.tocsr()))
    Lc = O.coarse_factor(Ac.toarray())
    dinv0 = O.jacobi_inv(O.ops_diagonal(ops, p.ndof), free)
    dd = torch.as_tensor(dinv0, device="cuda")
    A0 = O.op_bc(lambda w: O.ops_matvec(ops, w, p.ndof), free)
    rng = np.random.default_rng(11)
    Rm = free[None, :] * rng.standard_normal((5, p.ndof))
    Z = _np(amg.apply_multi(Rm, A0=BcOperator(op), dinv0=dd))
    for k in range(5):
        r = Rm[k]
        z0 = free * O.vcycle(lv, Lc, free * r, A0=A0, dinv0=dinv0) + (1 - free) * r
        assert rel(_np(amg.apply(r, A0=BcOperator(op), dinv0=dd)), z0) <= 1e-10, k
        assert rel(Z[k], z0) <= 1e-10, k
    assert rel(_np(amg.apply(Rm[0])), O.vcycle(lv, Lc, Rm[0])) <= 1e-10
    b = free * p.f if np.abs(p.f).sum() > 0 else free * rng.standard_normal(p.ndof)
    ud = spla.spsolve(A.tocsc(), b)
    for cycle in (L.F64, L.F32):
        M = AMGPrecond(amg, BcOperator(op), dinv0=dd, cycle=cycle)
        u, it, rr, st = pcg_device(BcOperator(op), b, M, tol=1e-10, maxiter=500)
        assert st == 0 and rel(_np(u), ud) <= 1e-8, (cycle, it, rel(_np(u), ud))


@pytest.mark.parametrize("kind,max_coarse,nc", [("tet10", 40, 72), ("hex", 40, 120), ("hex2", 100, 540)])
def test_device_coarse_factor_and_inverse_over_several_panels(gpu, kind, max_coarse, nc):
    """The device coarse solve (blocked Cholesky and explicit inverse in 32-row panels) on two-level
    hierarchies whose coarse matrices span 3, 4 and 17 panels with a ragged last one: the factor
    equals numpy's Cholesky of the shifted coarse matrix (gfem/linalg.py:126-133) and the V-cycle,
    whose coarse step multiplies by the inverse, equals the oracle's cho_solve cycle."""
    import torch
    from exfem_amd.linalg import AMG
    p, op, A, ops, free = _problem(kind)
    amg = AMG.from_operator(op, p.X, max_coarse=max_coarse, keep_levels=True)
    ref, Ac = O.sa_hierarchy(A, O.rigid_body_modes(p.X), 3, max_coarse=max_coarse)
    assert amg.sizes == [p.ndof, nc] and Ac.shape == (nc, nc)
    Lr = O.coarse_factor(Ac.toarray())
    assert np.allclose(np.triu(amg.Lc, 1), 0.0)
    assert mrel(amg.Lc, Lr) <= 1e-10
    assert mrel(amg.Lc @ amg.Lc.T, Lr @ Lr.T) <= 1e-12
    lv = []
    for r in ref:
        d = r["A"].diagonal()
        dinv = np.where(np.abs(d) > 0, 1.0 / np.where(d == 0, 1, d), 0.0)
        lv.append(dict(A=r["A"], dinv=dinv, rho=O.rho_estimate(r["A"], dinv), P=r["P"], R=r["P"].T.tocsr()))
    rng = np.random.default_rng(3)
    for k in range(3):
        r = free * rng.standard_normal(p.ndof)
        assert rel(_np(amg.apply(torch.as_tensor(r, device="cuda"))), O.vcycle(lv, Lr, r)) <= 1e-10, k


def test_device_sa_at_full_size_matches_the_host_builder(gpu):
    """The 1.01M-dof jittered C3D8 box (200x40x40): same level sizes, aggregates and rho as the
    host builder, P and the level-1 matrix to 1e-12, and a V-cycle equal to 1e-10."""
    import torch
    from exfem_amd import _lib as L
    from exfem_amd import workloads
    from exfem_amd.linalg import AMG, BcOperator
    p = workloads.jittered_cantilever(200, 40, 40)
    m = p.model()
    op = m.operator()
    op.set_ke(0, m.element_stiffness(storage=L.KE_SYM)[0])
    op.set_bc(p.free)
    dev = AMG.from_operator(op, p.X, keep_levels=True)
    host = AMG(op.assemble(bc=True), p.X)
    assert dev.sizes == host.sizes and len(dev.sizes) >= 3
    for lv, (d, h) in enumerate(zip(dev.levels_host, host.levels_host)):
        assert np.array_equal(d["agg"], h["agg"]), lv
        Pd, Ph = d["P"].tocsr(), h["P"].tocsr()
        assert abs(Pd - Ph).max() <= 1e-12 * abs(Ph).max(), lv
        if lv:
            Ad, Ah = sp.csr_matrix(d["A"]), sp.csr_matrix(h["A"])
            assert abs(Ad - Ah).max() <= 1e-12 * abs(Ah).max(), lv
        assert abs(d["rho"] - h["rho"]) <= 1e-12 * h["rho"], lv
    r = torch.as_tensor(p.free * np.random.default_rng(1).standard_normal(p.ndof), device="cuda")
    dinv0 = torch.as_tensor(host.levels_host[0]["dinv"], device="cuda")
    z1 = _np(dev.apply(r, A0=BcOperator(op), dinv0=dinv0))
    z2 = _np(host.apply(r, A0=BcOperator(op), dinv0=dinv0))
    assert rel(z1, z2) <= 1e-10
