"""Modal analysis on the device (gfem/eigen.py, gfem/solvers.py:1011-1071).

`lobpcg` is the block LOBPCG of gfem/eigen.py:32-80 for A x = lambda B x (smallest eigenvalues)
with blocks stored as rows, (m, ndof), the case-major layout of the batched product. `modal`
solves K phi = w^2 M phi with the HRZ lumped mass:
* eigensolver "lobpcg": A is the batched EbE product with the boundary mask, B the lumped mass on
  the free dofs, and the preconditioner the SA-AMG V-cycle (level 0 on the EbE operator) applied
  to every row (gfem/eigen.py:117-131). The Gram matrices are formed on the device; their small
  (at most 3m x 3m) FP64 eigenproblems are solved on the host, where a 45 x 45 eigh takes
  microseconds instead of the device solver's per-call overhead;
* eigensolver "lanczos": shift-invert eigsh of scipy on the device-assembled matrix (the
  reference's host branch, gfem/solvers.py:1028-1046).
"""
import time

import numpy as np

from . import _lib as L
from .model import as_dev, torch


def _gram(A, B, c=1024):
    """A B^T for row blocks A (m1, n), B (m2, n) with n large: batched products over 1024-column
    slices, then a sum over the slices (fixed order). The library GEMM runs this tall-skinny
    shape at about 25 GFLOP/s (162 ms for 45 x 45 over 1M columns); split this way it takes
    0.6 ms."""
    t = torch()
    n = A.shape[1]
    nc = n // c
    G = t.zeros(A.shape[0], B.shape[0], dtype=A.dtype, device=A.device)
    if nc:
        Ab = A[:, :nc * c].reshape(A.shape[0], nc, c).transpose(0, 1).contiguous()
        Bb = B[:, :nc * c].reshape(B.shape[0], nc, c).permute(1, 2, 0).contiguous()
        G = t.bmm(Ab, Bb).sum(0)
    if nc * c < n:
        G = G + A[:, nc * c:] @ B[:, nc * c:].T.contiguous()
    return G


def _eigh(G):
    """Eigen-decomposition of a small symmetric device matrix, solved on the host."""
    t = torch()
    w, V = np.linalg.eigh(G.cpu().numpy())
    return t.as_tensor(w, device=G.device), t.as_tensor(V, device=G.device)


def _svqb(Bop, S):
    """B-Gram matrix of the rows of S, scaled to unit diagonal, and its eigen-decomposition
    (gfem/eigen.py:21-29)."""
    t = torch()
    G = _gram(S, Bop(S))
    G = 0.5 * (G + G.T)
    d = t.sqrt(t.clamp(t.diagonal(G), min=1e-300))
    w, V = _eigh(G / d[:, None] / d[None, :])
    return w, V, d


def lobpcg(A, B, T, X0, k, tol=1e-6, maxiter=500, log=None):
    """gfem/eigen.py:32-80. A, B, T: functions on (m, n) blocks of row vectors; X0 (m, n) with
    m >= k. Returns (eigenvalues (k,) numpy, eigenvectors (k, n) device, iterations)."""
    t = torch()

    def orth(S):
        w, V, d = _svqb(B, S)
        keep = w > 1e-12 * w.max()
        Vk = V[:, keep] / t.sqrt(w[keep])[None, :]
        return Vk.T @ (S / d[:, None])

    def rr(S, m):
        AS = A(S)
        Ga = _gram(S, AS)
        Ga = 0.5 * (Ga + Ga.T)
        lam, C = _eigh(Ga)
        return lam[:m], C[:, :m], AS

    X = orth(X0)
    m = X.shape[0]
    lam, C, AS = rr(X, m)
    X = C.T @ X
    AX = C.T @ AS
    BX = B(X)
    P = None
    it = 0
    for it in range(1, maxiter + 1):
        R = AX - BX * lam[:, None]
        rn = t.linalg.norm(R, dim=1) / t.clamp(t.abs(lam) * t.linalg.norm(BX, dim=1), min=1e-300)
        res = float(rn[:k].max())
        if log and it % 10 == 1:
            log(f"    LOBPCG it {it:4d}: max rel. residual {res:.2e}")
        if res < tol:
            break
        W = T(R)
        W = W - _gram(W, BX) @ X                         # B-orthogonal to X
        parts = [X, W] if P is None else [X, W, P - _gram(P, BX) @ X]
        S = orth(t.cat(parts, 0))
        lam, C, AS = rr(S, m)
        Xn = C.T @ S
        P = Xn - _gram(Xn, BX) @ X                       # the update's part outside the old X
        X = Xn
        AX = C.T @ AS
        BX = B(X)
    return lam[:k].cpu().numpy(), X[:k], it


def _system(model, step, state):
    """Device operator with the step's boundary mask, attached FP64 element matrices, the free
    mask and the lumped mass."""
    from .solvers import _Setup, State
    st = state or State(model)
    setup = _Setup(model, step, st)
    op = model.operator()
    for i, K in enumerate(model.element_stiffness(storage=L.KE_SYM)):
        op.set_ke(i, K)
    op.set_bc(setup.free_u8)
    return op, setup, model.lumped_mass()


def modal_lobpcg(model, step, state=None, monitor=None, tol=1e-6):
    """gfem/eigen.py:98-117 on the device: (w2 (k,), modes (k, ndof) numpy, iterations)."""
    from .linalg import AMG, AMGPrecond, BcOperator
    from .solvers import Monitor
    t = torch()
    mon = monitor or Monitor(False)
    t0 = time.time()
    op, setup, M = _system(model, step, state)
    free = as_dev(setup.free_np)
    Mf = M * free
    d = op.diagonal()
    pinv = t.where(free > 0, 1.0 / t.where(d.abs() > 0, d, t.ones_like(d)), t.ones_like(d))
    t1 = time.time()
    amg = AMG.from_operator(op, model.X)
    pre = AMGPrecond(amg, BcOperator(op), pinv)
    t2 = time.time()
    mon.log(f"  modal (device LOBPCG): {int(setup.free_np.sum())} free dofs, {amg.info()}; "
            f"element matrices and mass {t1 - t0:.2f}s, AMG setup {t2 - t1:.2f}s")
    k = step.nmodes
    m = k + max(4, k // 2)
    # the blocks' products (m, 2m, 3m rows) on the assembled operator the AMG set-up already
    # built: no autotuning (and no patch plans) inside the solve
    for r in sorted({min(m, 64), min(2 * m, 64), min(3 * m, 64)}):
        op.set_multi_shape(r, 8)
    g = t.Generator().manual_seed(0)
    X0 = t.randn(m, model.ndof, generator=g, dtype=t.float64).to(free.device) * free[None, :]
    A = lambda V: t.cat([op.matvec_multi(V[i:i + 64].contiguous(), bc=True) for i in range(0, V.shape[0], 64)])
    B = lambda V: V * Mf[None, :]
    T = pre.multi                                          # batched V-cycles over the block
    lam, V, it = lobpcg(A, B, T, X0, k, tol=tol, log=mon.log)
    t.cuda.synchronize()
    mon.log(f"  LOBPCG converged in {it} iterations ({time.time() - t2:.2f}s)")
    return np.maximum(lam, 0.0), V.cpu().numpy(), it


def modal(model, step, state=None, monitor=None):
    """Natural frequencies (*FREQUENCY, gfem/solvers.py:1011-1071): K phi = w^2 M phi with the
    lumped mass. Returns dict(freq, modes, participation, effective_mass, total_mass) and, for
    LOBPCG, iterations."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from .solvers import Monitor
    mon = monitor or Monitor(False)
    t0 = time.time()
    method = getattr(step, "eigensolver", "auto") or "auto"
    if method == "auto":
        method = "lobpcg" if model.ndof > 30000 else "lanczos"
    extra = {}
    if method == "lobpcg":
        w2, modes, it = modal_lobpcg(model, step, state, mon)
        extra["iterations"] = it
        nm = len(w2)
    elif method == "lanczos":
        op, setup, Md = _system(model, step, state)
        keep = np.nonzero(setup.free_np > 0)[0]
        K = op.assemble().tocsr()[keep][:, keep].tocsc()
        M = Md.cpu().numpy()
        Mk = sp.diags(M[keep]).tocsc()
        if K.shape[0] <= max(step.nmodes + 2, 60):           # tiny systems: dense
            import scipy.linalg as sla
            nm = min(step.nmodes, K.shape[0])
            w2, V = sla.eigh(K.toarray(), Mk.toarray())
            w2, V = w2[:nm], V[:, :nm]
        else:
            nm = min(step.nmodes, K.shape[0] - 2)
            sigma = -abs(step.shift) if step.shift else -1e-6 * abs(K.diagonal()).max() / max(M.max(), 1e-300)
            w2, V = spla.eigsh(K, k=nm, M=Mk, sigma=sigma, which="LM")
        order = np.argsort(w2)
        w2, V = np.maximum(w2[order], 0.0), V[:, order]
        modes = np.zeros((nm, model.ndof))
        modes[:, keep] = V.T
    else:
        raise ValueError(f"unknown eigensolver {method!r}")
    freq = np.sqrt(w2) / (2 * np.pi)
    # participation factors and effective masses in x, y, z (gfem/solvers.py:1052-1062)
    Mfull = model.lumped_mass().cpu().numpy()
    r = np.zeros((3, model.ndof))
    for dd in range(3):
        r[dd, dd::3] = 1.0
    gm = np.einsum("mi,i,mi->m", modes, Mfull, modes)
    part = np.einsum("mi,i,di->md", modes, Mfull, r) / gm[:, None]
    meff = part ** 2 * gm[:, None]
    total = np.array([np.sum(Mfull[dd::3]) for dd in range(3)])
    for i, f in enumerate(freq):
        mon.log(f"    mode {i + 1:3d}: f = {f:.6g}  eff.mass ratio x/y/z = "
                + " ".join(f"{x:.3f}" for x in meff[i] / np.maximum(total, 1e-300)))
    mon.log(f"  modal done: {time.time() - t0:.2f}s")
    return dict(freq=freq, modes=modes, participation=part, effective_mass=meff, total_mass=total,
                **extra)


def _buckle_ops(model, K0ops, KGops, setup):
    """Device operators of K0 and KG (element matrices of the element blocks, FP64) with the
    step's boundary mask and MPC transformation."""
    from .model import EbeOperator, KeBlock
    op0 = model.operator()
    opG = EbeOperator(model.nn, [b.etype for b in model.blocks], [b.conn for b in model.blocks])
    for op, ops in ((op0, K0ops), (opG, KGops)):
        for i, (_, K) in enumerate(ops[:len(model.blocks)]):
            op.set_ke(i, K if isinstance(K, KeBlock) else KeBlock.from_array(model.blocks[i].etype, K))
        op.set_assembled(False)
        op.set_bc(setup.free_u8)
        setup.tr.bind(op)
    return op0, opG


def buckling_lobpcg(model, K0ops, KGops, setup, nmodes, monitor=None, tol=1e-6):
    """gfem/eigen.py:137-151 on the device: the most negative mu of KG phi = mu K0 phi by LOBPCG
    with the SA-AMG V-cycle of K0 as preconditioner; load factors -1/mu (inf for mu >= 0).
    Returns (load factors (k,) numpy, modes (k, ndof) numpy, iterations)."""
    from .linalg import AMG, AMGPrecond, BcOperator
    from .solvers import Monitor
    t = torch()
    mon = monitor or Monitor(False)
    if len(K0ops) > len(model.blocks) or setup.tr.active:
        raise NotImplementedError("device LOBPCG buckling takes element blocks without springs or MPCs: "
                                  "use eigensolver='lanczos'")
    op0, opG = _buckle_ops(model, K0ops, KGops, setup)
    free = as_dev(setup.free_np)
    d = op0.diagonal()
    pinv = t.where(free > 0, 1.0 / t.where(d.abs() > 0, d, t.ones_like(d)), t.ones_like(d))
    amg = AMG.from_operator(op0, model.X)
    pre = AMGPrecond(amg, BcOperator(op0), pinv)
    mon.log(f"  buckling (device LOBPCG): {int(setup.free_np.sum())} free dofs, {amg.info()}")
    m = nmodes + max(4, nmodes // 2)
    for op in (op0, opG):
        for r in sorted({min(m, 64), min(2 * m, 64), min(3 * m, 64)}):
            op.set_multi_shape(r, 8)
    g = t.Generator().manual_seed(1)
    X0 = t.randn(m, model.ndof, generator=g, dtype=t.float64).to(free.device) * free[None, :]
    mv = lambda op: (lambda V: t.cat([op.matvec_multi(V[i:i + 64].contiguous(), bc=True) * free[None, :]
                                      for i in range(0, V.shape[0], 64)]))
    mu, V, it = lobpcg(mv(opG), mv(op0), pre.multi, X0, nmodes, tol=tol, log=mon.log)
    lam = np.where(mu < 0, -1.0 / np.where(mu < 0, mu, -1.0), np.inf)
    mon.log(f"  LOBPCG converged in {it} iterations")
    return lam, V.cpu().numpy(), it


def buckling(model, step, state=None, monitor=None):
    """Linear (eigenvalue) buckling (*BUCKLE, gfem/solvers.py:1074-1117): (K0 + lambda KG(P)) phi
    = 0 with KG the initial-stress stiffness of the linear response to the step's loads P.
    eigensolver "lobpcg" (device, buckling_lobpcg) or "lanczos" (scipy eigsh on the device-assembled
    matrices, shift-inverted on K0); "auto" picks LOBPCG above 30,000 dof. Returns
    dict(load_factors, modes, base_u) - factors ascending."""
    import scipy.linalg as sla
    import scipy.sparse.linalg as spla
    from .linalg import bc_matrix
    from .solvers import Monitor, StaticSolver, State, Step, _Setup
    mon = monitor or Monitor(False)
    t0 = time.time()
    t = torch()
    st = state or State(model)
    solver = StaticSolver(model, linear_solver=getattr(step, "linear_solver", "auto") or "auto", verbose=False)
    s1 = Step("base", bcs=[(b[0], b[1], 0.0) for b in step.bcs], loads=step.loads, dloads=step.dloads)
    u, states, _ = solver.solve(s1, state=State(model))
    stress = model.stresses(u, states)
    K0ops = model.stiffness_ops(t.zeros(model.ndof, dtype=t.float64, device=u.device), model.init_state())
    KGops = model.geometric_ops(stress)
    setup = _Setup(model, step, st)
    method = getattr(step, "eigensolver", "auto") or "auto"
    if method == "auto":
        method = "lobpcg" if model.ndof > 30000 else "lanczos"
    extra = {}
    if method == "lobpcg":
        lam, modes, it = buckling_lobpcg(model, K0ops, KGops, setup, step.nmodes, mon)
        extra["iterations"] = it
    elif method == "lanczos":
        op0, opG = _buckle_ops(model, K0ops, KGops, setup)
        A0, AG = op0.assemble(), opG.assemble()
        if len(K0ops) > len(model.blocks):               # springs: summed on the host
            from .model import ops_assemble
            A0 = (A0.tocsr() + ops_assemble(K0ops[len(model.blocks):], model.ndof)).tocsr()
        if setup.tr.active:
            A0, AG = setup.tr.project_matrix(A0), setup.tr.project_matrix(AG)
        keep = np.nonzero(setup.free_np > 0)[0]
        K0 = A0.tocsr()[keep][:, keep].tocsc()
        KG = AG.tocsr()[keep][:, keep].tocsc()
        if K0.shape[0] <= max(step.nmodes + 2, 60):          # tiny systems: dense
            mu, V = sla.eigh(-KG.toarray(), K0.toarray())
        else:
            nm = min(step.nmodes, K0.shape[0] - 2)
            lu = spla.splu(K0)
            Kinv = spla.LinearOperator(K0.shape, matvec=lu.solve, dtype=float)
            mon.log(f"  buckling: {K0.shape[0]} free dofs, {nm} modes")
            mu, V = spla.eigsh(-KG, k=nm, M=K0, Minv=Kinv, which="LA")
        order = np.argsort(-mu)[:step.nmodes]
        mu, V = mu[order], V[:, order]
        lam = np.where(mu > 0, 1.0 / np.where(mu > 0, mu, 1.0), np.inf)
        modes = np.zeros((len(mu), model.ndof))
        modes[:, keep] = V.T
        if setup.tr.active:
            Tm = setup.tr.T
            modes = np.asarray([Tm @ mm for mm in modes])
    else:
        raise ValueError(f"unknown eigensolver {method!r}")
    order = np.argsort(lam)
    lam, modes = lam[order], modes[order]
    for i, l in enumerate(lam):
        mon.log(f"    mode {i + 1:3d}: load multiplier = {l:.6g}")
    mon.log(f"  buckling done: {time.time() - t0:.2f}s")
    return dict(load_factors=lam, modes=modes, base_u=u, **extra)


__all__ = ["lobpcg", "modal_lobpcg", "modal", "buckling_lobpcg", "buckling"]
