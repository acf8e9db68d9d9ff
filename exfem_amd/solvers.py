"""Linear-system dispatch and the static procedure on the MI355X path
(gfem/solvers.py:24-95,104-231,235-477,481-683; linear-elastic subset).

``_LinearSystem`` keeps the reference's constructor and ``solve(ops, b, setup, jvp_args,
rebuild)`` signature and back-end names (auto | direct | amg | jacobi | cheby | mixed); every
iterative method runs on the device through the HIP library. One back-end is added: ``amg32``,
FP64 AMG-PCG whose V-cycle runs in FP32 (FP64 operator, residual, dot products, norms and
coarse solve), the same iteration count as ``amg`` at a cheaper cycle.
"""
import concurrent.futures as cf
import time

import numpy as np

from . import _lib as L
from .linalg import (AMG, AMGPrecond, BcOperator, Chebyshev, Direct, Jacobi, bc_matrix,
                     lmax_power, pcg_device)
from .model import KeBlock, as_dev, ops_assemble, set_rest_ops, to_numpy, torch


class Monitor:
    """gfem/solvers.py:24-45."""

    def __init__(self, verbose=True):
        self.verbose = verbose
        self.stop = False

    def log(self, msg):
        if self.verbose:
            print(msg, flush=True)

    def iteration(self, **kw):
        pass

    def increment(self, **kw):
        pass

    def frame(self, t, u, states, **kw):
        pass

    def should_stop(self):
        return self.stop


class Step:
    """gfem/solvers.py:48-80 (static and frequency subset): bcs [(nodes, dof, value)], loads
    [(nodes, dof, value)], dloads [(vector,)]; nmodes, shift and eigensolver for modal()."""

    def __init__(self, name="Step-1", kind="static", time_period=1.0, dt_init=None, dt_min=None,
                 dt_max=None, max_inc=1000, bcs=(), loads=(), pressures=(), dloads=(),
                 temperature=None, nmodes=10, shift=0.0, eigensolver="auto", **kw):
        self.nmodes, self.shift, self.eigensolver = int(nmodes), float(shift), eigensolver
        self.name, self.kind, self.T = name, kind, float(time_period)
        self.dt_init = dt_init or self.T
        self.dt_min = dt_min if dt_min is not None else 1e-5 * self.T
        self.dt_max = dt_max or self.T
        self.max_inc = max_inc
        self.bcs, self.loads = list(bcs), list(loads)
        self.pressures, self.dloads = list(pressures), list(dloads)
        self.temperature = temperature
        if self.pressures:
            raise NotImplementedError("pressure loads: pass them as dloads vectors")


class State:
    """gfem/solvers.py:83-95."""

    def __init__(self, model):
        self.u = np.zeros(model.ndof)
        self.states = model.init_state()
        self.fext = np.zeros(model.ndof)
        self.T = None                  # nodal temperature at the end of the last step
        self.time = 0.0
        self.reactions = np.zeros(model.ndof)


def _ramp(t, T):
    return min(max(t / T, 0.0), 1.0) if T > 0 else 1.0


def _temperature(setup, t):
    """_Setup.temperature (gfem/solvers.py:215-217) with the default linear ramp."""
    r = _ramp(t, setup.step.T)
    return setup.Tprev + (setup.Tnew - setup.Tprev) * r


class _Setup:
    """Host BCs and dead loads of one step (gfem/solvers.py:104-217, no MPC / amplitudes)."""

    def __init__(self, model, step, state=None):
        m = model
        nd = m.ndof
        self.step = step
        self.fixed = np.zeros(nd, bool)
        self.target = np.zeros(nd)
        for bc in step.bcs:
            nodes, dof, val = bc[:3]
            d = np.asarray(nodes, dtype=np.int64).ravel() * 3 + dof
            self.fixed[d] = True
            self.target[d] = val
        g = np.zeros(nd)
        for ld in step.loads:
            nodes, dof, val = ld[:3]
            np.add.at(g, np.asarray(nodes, dtype=np.int64).ravel() * 3 + dof, val)
        for dl in step.dloads:
            g += np.asarray(dl[0] if isinstance(dl, (tuple, list)) else dl, float)
        self.load = g
        self.fprev = np.zeros(nd) if state is None else np.asarray(state.fext)
        self.ustart = np.zeros(nd) if state is None else np.asarray(to_numpy(state.u))
        # temperatures (gfem/solvers.py:159-162,215-217): from the last step's end, else the
        # step's own field, else the model's initial temperature; ramped to the step's field
        T0 = np.zeros(m.nn) if getattr(m, "T0", None) is None else np.asarray(m.T0, float)
        Tst = None if getattr(step, "temperature", None) is None else \
            np.broadcast_to(np.asarray(step.temperature, float), (m.nn,)).copy()
        Tst_prev = None if state is None else getattr(state, "T", None)
        self.Tprev = Tst_prev if Tst_prev is not None else (Tst if Tst is not None else T0)
        self.Tnew = Tst if Tst is not None else self.Tprev
        self.thermal = Tst is not None or getattr(m, "T0", None) is not None
        # MPC elimination (gfem/solvers.py:164-167): slaves leave the free set
        from .constraints import Transformation
        self.tr = Transformation(nd, m.constraints, self.fixed)
        if self.tr.dropped:
            print(f"  warning: {self.tr.dropped} constrained dofs are also prescribed (constraint dropped)")
        self.free_np = (~self.fixed & ~self.tr.slave_mask).astype(float)
        self.free_u8 = (~self.fixed & ~self.tr.slave_mask).astype(np.uint8)
        self.tr.bind(m.operator())

    def fext(self, t):
        r = _ramp(t, self.step.T)
        return self.fprev * (1 - r) + self.load * r

    def prescribed(self, t):
        r = _ramp(t, self.step.T)
        tgt = self.ustart + (self.target - self.ustart) * r
        return np.where(self.fixed, tgt, 0.0)


class _LinearSystem:
    """Solves (free K free + fixed I) x = b (gfem/solvers.py:235-477)."""

    def __init__(self, model, method="auto", tangent="ebe", cg_tol=1e-9, cg_maxiter=20000,
                 monitor=None, multi_gpu=True, operator="auto", ke_storage=L.KE_FULL):
        self.m = model
        if method == "auto":
            method = "direct" if model.ndof <= 20000 else "amg"
        if method not in ("direct", "amg", "amg32", "jacobi", "cheby", "mixed", "none"):
            raise ValueError(f"unknown linear solver {method!r}")
        self.method, self.tangent = method, tangent
        if operator not in ("auto", "gather", "csr"):
            raise ValueError(f"unknown operator {operator!r}")
        self.operator = "gather" if operator == "auto" else operator
        self.async_amg = 150000
        self.cg_tol, self.cg_maxiter = cg_tol, cg_maxiter
        self.ke_storage = ke_storage
        self.amg = None
        self.amg_iters0 = None
        self.mon = monitor or Monitor(False)
        self.last_iters = 0
        self.last_relres = 0.0
        self._pool = None
        self._amg_future = None

    # ------------------------------------------------------------------ operator plumbing
    def _split(self, ops):
        """(element-block ops, rest ops): the model's element blocks come first
        (gfem/solvers.py:401-405), unless ops holds only rest operators (the jvp tangent
        without element matrices, gfem/solvers.py:500-504)."""
        nb = len(self.m.blocks)
        if len(ops) >= nb and all(isinstance(K, KeBlock) or (np.shape(K)[-1] // 3 in (8, 10))
                                  for _, K in ops[:nb]):
            return ops[:nb], ops[nb:]
        return [], list(ops)

    def _device_op(self, ops, setup, dtype=L.F64, fields=None, matrix_free=False):
        # the matrix-free products read the geometry and the solve's fields (the per-element
        # scale); stored products read neither, so the operator's geometry is left as it is
        op = self.m.operator(fields) if matrix_free else self.m.live_operator()
        blocks, rest = self._split(ops)
        for i, (_, K) in enumerate(blocks):
            k = K if isinstance(K, KeBlock) else KeBlock.from_array(self.m.blocks[i].etype, K)
            if k.dtype != dtype:
                k = k.astype(dtype)
            op.set_ke(i, k)
        set_rest_ops(op, rest)          # springs, contact pairs: the rest of opA
        # operator="csr" (gfem/solvers.py:406-414): the element blocks' products on their
        # assembled matrix for the Jacobi and AMG solves; the rest stays element by element
        op.set_assembled(self.operator == "csr" and self.method in ("amg", "amg32", "jacobi")
                         and bool(blocks) and not matrix_free)
        op.set_bc(setup.free_u8)
        setup.tr.bind(op)
        return op

    def _pinv(self, op, setup, have_ke=True):
        free = as_dev(setup.free_np)
        if not have_ke:
            return torch().where(free > 0, torch().ones_like(free), torch().ones_like(free))
        d = op.diagonal()
        safe = torch().where(d.abs() > 0, d, torch().ones_like(d))
        return torch().where(free > 0, 1.0 / safe, torch().ones_like(d))

    def _lmax(self, A, dinv, setup):
        """25 power iterations of D^-1 A from the reference's sine start (gfem/solvers.py:328-342),
        in the library (exfem_pcg_lmax); the start's free mask is the operator's boundary mask,
        which set_bc took from setup.free."""
        return lmax_power(A, dinv, 25)

    def assemble(self, ops, setup):
        """bc_matrix(ops_assemble(ops)) (gfem/solvers.py:323-326); the element sums run on the
        device when `ops` are this model's element blocks."""
        blocks, rest = self._split(ops)
        if blocks and len(blocks) == len(self.m.blocks) and all(isinstance(K, KeBlock) for _, K in blocks):
            op = self._device_op(ops, setup, L.F64)
            if not setup.tr.active and not rest:
                return op.assemble(bc=True)
            A = op.assemble(bc=False)
            if rest:                     # springs, contact pairs: summed on the host
                A = (A.tocsr() + ops_assemble(rest, self.m.ndof)).tocsr()
        else:
            A = ops_assemble(ops, self.m.ndof)
        return bc_matrix(setup.tr.project_matrix(A), setup.free_np)

    def _device_amg_ok(self, ops, setup):
        """The SA set-up runs on the device (AMG.from_operator) when A is the element blocks'
        assembled matrix with the boundary mask: no rest operators, no MPC projection, nodal
        coordinates for the rigid-body candidates."""
        blocks, rest = self._split(ops)
        return (bool(blocks) and len(blocks) == len(self.m.blocks) and not rest
                and all(isinstance(K, KeBlock) for _, K in blocks) and not setup.tr.active
                and self.m.X is not None and 3 * len(self.m.X) == self.m.ndof)

    def _build_amg(self, ops, setup):
        t0 = time.time()
        if self._device_amg_ok(ops, setup):
            amg = AMG.from_operator(self._device_op(ops, setup, L.F64), self.m.X)
        else:
            amg = AMG(self.assemble(ops, setup), self.m.X)
        self.mon.log(f"    {amg.info()}  setup {time.time() - t0:.2f}s")
        return amg

    # ------------------------------------------------------------------ back-ends
    def _run(self, A, b, M, tol=None, maxiter=None):
        # a background AMG build uploads its levels with synchronous copies, which must not
        # overlap a stream capture: while one is in flight the iterations launch eagerly
        from . import linalg
        busy = self._amg_future is not None and not self._amg_future.done()
        prev = linalg._USE_GRAPH[0]
        if busy:
            linalg.pcg_graphs(False)
        try:
            x, k, rr, st = pcg_device(A, b, M, self.cg_tol if tol is None else tol,
                                      self.cg_maxiter if maxiter is None else maxiter)
        finally:
            if busy:
                linalg.pcg_graphs(prev)
        return x, k, rr

    def _mixed(self, ops, b, setup, rebuild):
        """FP32 AMG-PCG inner solves (tol 2e-5) inside <= 12 FP64 refinement steps
        (gfem/solvers.py:344-384). FP64: the residual b - A x, the refinement update, all PCG
        dot products and norms, the coarse solve."""
        if self.amg is None or rebuild:
            self.amg = self._build_amg(ops, setup)
        self.amg.set_precision(L.F32)
        op64 = self._device_op(ops, setup, L.F64)
        pinv = self._pinv(op64, setup)
        A64 = BcOperator(op64)
        A32 = BcOperator(op64, dtype=L.F32)
        for i, (_, K) in enumerate(ops[:len(self.m.blocks)]):
            k = K if isinstance(K, KeBlock) else KeBlock.from_array(self.m.blocks[i].etype, K)
            op64.set_ke(i, k.astype(L.F32, storage=k.storage))
        M32 = AMGPrecond(self.amg, A32, pinv, L.F32)
        b = as_dev(b)
        x = torch().zeros_like(b)
        r = b.clone()
        bn = float(torch().linalg.norm(b)) or 1.0
        its = 0
        for _ in range(12):
            sc = float(torch().linalg.norm(r)) or 1.0
            d, k, _ = self._run(A32, (r / sc).float(), M32, tol=2e-5)
            its += int(k)
            x = x + d.double() * sc
            r = b - A64(x)
            if float(torch().linalg.norm(r)) <= self.cg_tol * bn:
                break
        self.last_iters = its
        return x

    def solve(self, ops, b, setup, jvp_args=None, rebuild=False):
        if self.method == "direct":
            if jvp_args is not None:
                raise ValueError("direct solver requires tangent='ebe'")
            x = Direct(self.assemble(ops, setup)).solve(to_numpy(b))
            self.last_iters = 1
            return as_dev(x)
        if self.method == "mixed" and jvp_args is None:
            return self._mixed(ops, b, setup, rebuild)
        mf = jvp_args is not None
        fields = jvp_args[2] if mf and len(jvp_args) > 2 else None
        if mf:
            # products apply the tangent matrix-free (Model.tangent_jvp with the solve's fields,
            # gfem/solvers.py:280-283). The element matrices, when given, only build the
            # preconditioner (its diagonal or the AMG hierarchy); the reference passes none in
            # jvp mode (gfem/solvers.py:500-504), which StaticSolver(jvp_precond="none") follows.
            if ops:
                op = self._device_op(ops, setup, fields=fields, matrix_free=True)
            else:
                op = self.m.operator(fields)
                op.set_bc(setup.free_u8)
                setup.tr.bind(op)
            if self.m.nonlinear:               # J2 / hyperelastic blocks: the tangent at (u, states)
                u, st = jvp_args[0], jvp_args[1]
                op.set_state(u if u is not None else torch().zeros(self.m.ndof, dtype=torch().float64,
                                                                    device=as_dev(setup.free_np).device),
                             self.m.init_state() if st is None else st)
        else:
            op = self._device_op(ops, setup)
            if self.method == "amg32":         # FP32 element matrices for the FP32 level 0
                for i, (_, K) in enumerate(ops[:len(self.m.blocks)]):
                    k = K if isinstance(K, KeBlock) else KeBlock.from_array(self.m.blocks[i].etype, K)
                    op.set_ke(i, k.astype(L.F32, storage=k.storage))
        A = BcOperator(op, matrix_free=mf)
        pinv = self._pinv(op, setup, have_ke=bool(ops) or not mf)
        if self.method == "none":
            x, k, rr = self._run(A, b, None)
        elif self.method == "jacobi":
            x, k, rr = self._run(A, b, Jacobi(pinv))
        elif self.method == "cheby":
            x, k, rr = self._run(A, b, Chebyshev(pinv, self._lmax(A, pinv, setup), A))
        else:
            fut = self._amg_future
            if fut is not None and fut.done():
                self.amg, self._amg_future = fut.result(), None
                self.amg_iters0 = None
                rebuild = False            # a hierarchy that just landed is tried first
            if (self.amg is None or rebuild) and self._amg_future is None:
                # a host-built first hierarchy builds in the background while Chebyshev-Jacobi
                # PCG runs (gfem/solvers.py:443-453); the device set-up (0.14 s at 1M dof) and a
                # rebuild - the hierarchy in hand has slowed to 4x its first iteration count -
                # build at once, since every solve on the stale one until a background build
                # lands costs more than the build
                if self.m.ndof >= self.async_amg and not rebuild and not self._device_amg_ok(ops, setup):
                    A_host = self.assemble(ops, setup)
                    self._pool = self._pool or cf.ThreadPoolExecutor(1)
                    self._amg_future = self._pool.submit(AMG, A_host, self.m.X)
                    self.mon.log("    AMG setup started in background; Chebyshev-Jacobi PCG meanwhile")
                else:
                    self.amg = self._build_amg(ops, setup)
                    self.amg_iters0 = None
            if self.amg is None or (rebuild and self._amg_future is not None):
                x, k, rr = self._run(A, b, Chebyshev(pinv, self._lmax(A, pinv, setup), A))
                self.last_iters, self.last_relres = int(k), rr
                return x
            cyc = L.F32 if self.method == "amg32" else None
            if cyc is not None and mf and self.m.nonlinear:
                # the FP32 cycle's level 0 runs the FP32 matrix-free pass, which has no J2 or
                # hyperelastic tangent: such a jvp solve keeps the FP64 cycle
                if not getattr(self, "_warned32", False):
                    self.mon.log("    amg32 with a matrix-free nonlinear tangent: V-cycle in FP64")
                    self._warned32 = True
                cyc = None
            # the rebuild rule (k > 4 k0) is known once 4 k0 + 1 iterations have run: stop there
            # and rebuild rather than finish on the stale hierarchy
            k0 = self.amg_iters0
            cap = None if (k0 is None or rebuild) else min(4 * k0 + 1, self.cg_maxiter)
            x, k, rr = self._run(A, b, AMGPrecond(self.amg, A, pinv, L.F64, cycle=cyc), maxiter=cap)
            if self.amg_iters0 is None:
                self.amg_iters0 = max(k, 5)
            elif k > 4 * self.amg_iters0 and not rebuild:
                return self.solve(ops, b, setup, jvp_args, rebuild=True)
        self.last_iters, self.last_relres = int(k), rr
        if not np.isfinite(rr) or rr > 1e-3:
            self.mon.log(f"    warning: PCG not converged ({int(k)} it, rel.res {rr:.2e})")
        return x

    def wait_amg(self):
        """Block until a background AMG setup has finished (and adopt it)."""
        if self._amg_future is not None:
            self.amg, self._amg_future = self._amg_future.result(), None
            self.amg_iters0 = None
        return self.amg

    def reset(self):
        self.amg = None


class StaticSolver:
    """Implicit static step (gfem/solvers.py:481-683): tangent predictor from the converged
    state, Newton iterations with line search, states committed per converged increment,
    cutback to a quarter on failure and growth by 1.5 after an easy increment, reactions.
    Loads ramp linearly over the step. J2 blocks re-form their consistent tangent at every
    iteration; linear blocks keep theirs."""

    def __init__(self, model, rtol=5e-3, max_iter=16, linear_solver="auto", tangent="ebe",
                 cg_tol=1e-9, cg_maxiter=20000, verbose=True, monitor=None, line_search=True,
                 ke_storage=L.KE_FULL, jvp_precond="ke", operator="auto"):
        """jvp_precond (tangent='jvp' only): "ke" forms the tangent's element matrices at every
        Newton iteration to build the preconditioner (Jacobi diagonal, Chebyshev bound or AMG
        hierarchy) while every product stays matrix-free; "none" forms none, as the reference's
        jvp path does (gfem/solvers.py:500-504): Jacobi is then the identity, and AMG, which
        needs an assembled matrix, is refused."""
        if jvp_precond not in ("ke", "none"):
            raise ValueError(f"unknown jvp_precond {jvp_precond!r}")
        if tangent == "jvp" and jvp_precond == "none" and linear_solver in ("amg", "amg32", "auto", "mixed"):
            raise ValueError("AMG needs the tangent's element matrices: use jvp_precond='ke' "
                             "or linear_solver='jacobi' / 'cheby'")
        self.jvp_precond = jvp_precond
        self.m, self.rtol, self.max_iter = model, rtol, max_iter
        self.mon = monitor or Monitor(verbose)
        self.line_search = line_search
        self.lin = _LinearSystem(model, linear_solver, tangent, cg_tol, cg_maxiter, self.mon,
                                 ke_storage=ke_storage, operator=operator)
        self._ops = None
        self._ops_key = None
        self._lam = 1.0

    def _interactions(self, u, force=True, stiff=True):
        """[(conn, fe, K)] of the model's contacts and walls at u (gfem/solvers.py:494-504,510-517):
        parameters (host search, device selection) and element forces / tangents on the device,
        walls moved by the current load factor."""
        out = []
        for c in self.m.interactions:
            p = c.params(self.m.X, u, self._lam)
            fe, K = c.evaluate(p, force, stiff)
            out.append((c.conn(p), fe, K))
        return out

    def _fint(self, u, s_old, fields):
        """(f, trial states) at u from the converged states s_old; contact forces included."""
        f, s = self.m.internal_force(u, s_old, fields)
        if self.m.interactions:
            op = self.m.live_operator()
            base = len(self.m.discrete)
            for j, (conn, fe, _) in enumerate(self._interactions(u, stiff=False)):
                op.ensure_rest(base + j, conn)
                op.rest_assemble(base + j, fe, f)
        return f, s

    def _tangent_ops(self, u, s_old, fields):
        inter = []
        if self.m.interactions:
            inter = [((np.asarray(conn, np.int64)[:, :, None] * 3 + np.arange(3)).reshape(len(conn), -1), K)
                     for conn, _, K in self._interactions(u, force=False)]
        if self.lin.tangent == "jvp" and self.jvp_precond == "none":
            return inter
        if self.m.nonlinear or (self.m.discrete and any(b.kind == "axial" for b in self.m.discrete)):
            return self.m.stiffness_ops(u, s_old, fields, storage=self.lin.ke_storage) + inter
        key = None if not fields else id(fields.get("scale"))
        if self._ops is None or self._ops_key != key:
            self._ops = self.m.stiffness_ops(None, None, fields, storage=self.lin.ke_storage)
            self._ops_key = key
        return self._ops + inter

    def _linsolve(self, u, s_old, fields, b, setup):
        ops = self._tangent_ops(u, s_old, fields)
        jv = (u, s_old, fields) if self.lin.tangent == "jvp" else None
        return self.lin.solve(ops, b, setup, jvp_args=jv)

    def solve(self, step, u0=None, states=None, fext0=None, callback=None, state=None, scale=None):
        m = self.m
        st = state or State(m)
        if u0 is not None:
            st.u = u0
        if states is not None:
            st.states = states
        setup = _Setup(m, step, st)
        self.setup = setup
        self.lin.reset()
        fields = dict(scale=scale) if scale is not None else None
        fields_at = lambda tt: (dict(scale=scale, T=as_dev(_temperature(setup, tt))) if setup.thermal
                                else fields)
        free = as_dev(setup.free_np)
        fixed = 1.0 - free
        u = as_dev(st.u)
        states = st.states
        t, dt, inc, history = 0.0, min(step.dt_init, step.T), 0, []
        t0 = time.time()
        total_iters = 0
        self.qref = 0.0
        while t < step.T * (1 - 1e-12):
            if self.mon.should_stop():
                self.mon.log("  stopped by user")
                break
            inc += 1
            if inc > step.max_inc:
                raise RuntimeError("max increments exceeded")
            dt = min(dt, step.T - t)
            tn = t + dt
            self._lam = tn / step.T if step.T > 0 else 1.0     # load factor (wall motion)
            fext = as_dev(setup.fext(tn))
            fields = fields_at(tn)                     # the temperature at the increment's end
            ut = torch().where(fixed > 0, as_dev(setup.prescribed(tn)), u)
            ut = setup.tr.apply(ut) if setup.tr.active else ut
            ut = self._predictor(u, states, ut, fext, fields, setup, free)
            ok, it, u_new, s_try = self._newton(ut, states, fext, fields, setup, free)
            total_iters += it
            if ok:
                u, states, t = u_new, s_try, tn
                history.append((t, it))
                self.mon.log(f"  inc {inc:4d}  t={t:.4g}  dt={dt:.3g}  iters={it}  lin={self.lin.last_iters}")
                if callback:
                    callback(t, u, states)
                if it <= 5:
                    dt = min(dt * 1.5, step.dt_max)
            else:
                dt *= 0.25
                self.mon.log(f"  inc {inc:4d}  cutback -> dt={dt:.3g}")
                if dt < step.dt_min:
                    raise RuntimeError("time increment smaller than minimum")
        self.mon.log(f"  step '{step.name}' done: {inc} incs, {total_iters} iterations, "
                     f"{time.time() - t0:.2f}s")
        fint, _ = self._fint(u, states, fields_at(step.T))
        fe = as_dev(setup.fext(step.T))
        reactions = (fint - fe) * fixed
        st.u, st.states, st.fext = u, states, to_numpy(fe)
        st.T = setup.Tnew if setup.thermal else getattr(st, "T", None)
        st.time, st.reactions = st.time + step.T, reactions
        return u, states, dict(history=history, reactions=reactions, fext=fe, state=st)

    def _predictor(self, uc, s_old, ut, fext, fields, setup, free):
        """gfem/solvers.py:607-624: tangent predictor from the converged state."""
        delta = ut - uc
        fint, _ = self._fint(uc, s_old, fields)
        if self.lin.tangent == "jvp":
            Kd = self.m.tangent_jvp(uc, s_old, delta, fields)
        else:
            from .model import ops_matvec
            Kd = ops_matvec(self._tangent_ops(uc, s_old, fields), delta, self.m.ndof)
        r = free * setup.tr.apply_T(fint + Kd - fext)
        rn = float(torch().linalg.norm(r))
        self.qref = max(self.qref, rn)
        if rn == 0.0:
            return ut
        du = self._linsolve(uc, s_old, fields, -r, setup)
        u = ut + free * du
        return setup.tr.apply(u) if setup.tr.active else u

    def _newton(self, u, s_old, fext, fields, setup, free):
        """gfem/solvers.py:626-683 (no rotation dofs): (converged, iterations, u, states)."""
        fn = float(torch().linalg.norm(fext))
        dut = 0.0
        for it in range(1, self.max_iter + 1):
            fint, s_new = self._fint(u, s_old, fields)
            r = free * setup.tr.apply_T(fint - fext)
            rt = float(torch().linalg.norm(r))
            ft = float(torch().linalg.norm(fint))
            ref = max(fn, ft, 1e-3 * self.qref, 1e-30)
            self.qref = max(self.qref, fn, ft)
            ratio = rt / ref
            if not np.isfinite(rt):
                return False, it, u, s_old
            small_du = dut <= 1e-2 * max(float(torch().linalg.norm(u)), 1e-30)
            if ratio <= self.rtol * 1e-2 or (it > 1 and ratio <= self.rtol and small_du):
                return True, it - 1, u, s_new
            du = self._linsolve(u, s_old, fields, -r, setup)
            alpha = 1.0
            if self.line_search:
                for _ in range(4):
                    ut = u + alpha * du
                    ut = setup.tr.apply(ut) if setup.tr.active else ut
                    r2 = free * setup.tr.apply_T(self._fint(ut, s_old, fields)[0] - fext)
                    if float(torch().linalg.norm(r2)) < rt or alpha < 0.2:
                        break
                    alpha *= 0.5
            u = u + alpha * du
            u = setup.tr.apply(u) if setup.tr.active else u
            dut = float(torch().linalg.norm(alpha * du))
        return False, self.max_iter, u, s_old
