"""Linear multi-point constraints by master-slave elimination (gfem/constraints.py:1-25,198-302).

Every constraint is a row u[slave] = sum_k w_k u[master_k]. ``Transformation`` merges them into
u = T u_bar with the library's host build (``exfem_mpc_build``: one constraint per slave dof,
prescribed slaves dropped, chains resolved, cycles solved or broken) and hands T to the device
operator (``exfem_op_set_mpc``),
which then applies opA as free * T^T K T(free * v) + (1 - free) * v in every product with the
boundary mask. Tied contact (surface projection) is not on the device path.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .model import as_dev, stream_ptr, torch, vptr


class MPC:
    """Constraint rows u[slave_i] = sum_k weights[i, k] u[masters[i, k]]
    (gfem/constraints.py:14-25). masters / weights are (nc, k); a flat masters array means one
    master per row."""

    def __init__(self, slave, masters, weights, name="mpc"):
        s = np.asarray(slave, dtype=np.int64).reshape(-1)
        m = np.asarray(masters, dtype=np.int64)
        m = m.reshape(len(s), m.size // len(s) if len(s) else 0)
        self.slave, self.masters = s, m
        self.weights = np.asarray(weights, dtype=np.float64).reshape(m.shape)
        self.name = name

    def __len__(self):
        return len(self.slave)


def normal_constraint(nodes, normal, name="normal"):
    """n . u = 0 at every node (a frictionless support or a symmetry plane of any orientation,
    gfem/constraints.py:198-209). The component with the largest |n_j| is eliminated:
    u_k = -sum_{j != k} (n_j / n_k) u_j."""
    n = np.asarray(normal, dtype=np.float64).reshape(3)
    n = n / np.linalg.norm(n)
    k = int(np.argmax(np.abs(n)))
    rest = np.array([j for j in range(3) if j != k])
    nodes = np.asarray(nodes, dtype=np.int64).reshape(-1)
    return MPC(3 * nodes + k, 3 * nodes[:, None] + rest[None, :],
               np.broadcast_to(-n[rest] / n[k], (len(nodes), 2)), name)


def coupling(ref_dofs_nodes, nodes, name="coupling"):
    """Kinematic coupling of translations (gfem/constraints.py:212-218): the three displacement
    components of every node other than the reference node equal the reference node's."""
    ref = int(ref_dofs_nodes)
    nodes = np.asarray(nodes, dtype=np.int64).reshape(-1)
    nodes = nodes[nodes != ref]
    comp = np.arange(3)
    slave = (3 * nodes[:, None] + comp).reshape(-1)
    master = np.tile(3 * ref + comp, len(nodes)).reshape(-1, 1)
    return MPC(slave, master, np.ones(master.shape), name)


class Transformation:
    """u = T u_bar for the constraint rows of `mpcs` (gfem/constraints.py:221-302). T is built on
    the host by the library (exfem_mpc_build: first row per slave wins, prescribed slaves
    dropped, slave chains resolved through the strongly connected components of the slave graph,
    cycles taken to their fixed point or broken); products with T and T^T run on the device
    operator it is bound to (``bind``)."""

    def __init__(self, ndof, mpcs, fixed=None):
        self.ndof = int(ndof)
        groups = [m for m in mpcs if len(m)]
        self.active = bool(groups)
        self.slave_mask = np.zeros(self.ndof, bool)
        self.dropped = self.broken = 0
        self.T = None
        self._op = None
        if not self.active:
            return
        kmax = max(m.masters.shape[1] for m in groups)
        nrow = sum(len(m) for m in groups)
        slave = np.empty(nrow, np.int64)
        masters = np.full((nrow, max(kmax, 1)), -1, np.int64)
        weights = np.zeros((nrow, max(kmax, 1)))
        o = 0
        for m in groups:
            n, k = m.masters.shape
            slave[o:o + n] = m.slave
            masters[o:o + n, :k] = m.masters
            weights[o:o + n, :k] = m.weights
            o += n
        fx = None if fixed is None else np.ascontiguousarray(np.asarray(fixed, bool), np.uint8)
        lib = L.load()
        plan = C.c_void_p()
        L.check(lib.exfem_mpc_build(self.ndof, nrow, masters.shape[1], slave.ctypes.data,
                                    masters.ctypes.data, weights.ctypes.data,
                                    None if fx is None else fx.ctypes.data, C.byref(plan)))
        try:
            cnt = [C.c_int64() for _ in range(4)]
            L.check(lib.exfem_mpc_plan_info(plan, *[C.byref(c) for c in cnt]))
            nnz = cnt[0].value
            ptr = np.empty(self.ndof + 1, np.int64)
            cols = np.empty(max(nnz, 1), np.int32)
            vals = np.empty(max(nnz, 1))
            smask = np.empty(self.ndof, np.uint8)
            L.check(lib.exfem_mpc_plan_get(plan, ptr.ctypes.data, cols.ctypes.data,
                                           vals.ctypes.data, smask.ctypes.data))
        finally:
            lib.exfem_mpc_plan_free(plan)
        self.dropped, self.broken = cnt[2].value, cnt[3].value
        self.slave_mask = smask.astype(bool)
        self.T = sp.csr_matrix((vals[:nnz], cols[:nnz], ptr), shape=(self.ndof, self.ndof))

    def triplets(self):
        coo = self.T.tocoo()
        return (np.ascontiguousarray(coo.row, dtype=np.int32), np.ascontiguousarray(coo.col, dtype=np.int32),
                np.ascontiguousarray(coo.data, dtype=np.float64))

    def bind(self, op):
        """Attach T to a device operator (exfem_op_set_mpc; an inactive transformation clears
        it) and use that operator for apply / apply_T."""
        if not self.active:
            L.check(L.load().exfem_op_set_mpc(op.h, 0, None, None, None))
            self._op = None
            return self
        r, c, v = self.triplets()
        L.check(L.load().exfem_op_set_mpc(op.h, len(v), r.ctypes.data, c.ctypes.data, v.ctypes.data))
        self._op = op
        return self

    def _apply(self, v, transpose):
        vd = as_dev(v)
        y = torch().empty_like(vd)
        L.check(L.load().exfem_op_mpc_apply(self._op.h, int(transpose), vptr(vd), vptr(y), stream_ptr()))
        return y

    def apply(self, v):
        """T v (Transformation.apply, gfem/constraints.py:278-281)."""
        if not self.active:
            return v
        return self._apply(v, False)

    def apply_T(self, y):
        """T^T y (Transformation.apply_T, gfem/constraints.py:283-286)."""
        if not self.active:
            return y
        return self._apply(y, True)

    def project_matrix(self, A):
        """T^T A T for a scipy matrix (the direct and AMG set-up paths)."""
        if not self.active:
            return A
        return (self.T.T @ A @ self.T).tocsr()


__all__ = ["MPC", "normal_constraint", "coupling", "Transformation"]
