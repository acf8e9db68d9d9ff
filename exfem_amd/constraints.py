"""Linear multi-point constraints by master-slave elimination (gfem/constraints.py:1-25,198-302).

Every constraint is a row u[slave] = sum_k w_k u[master_k]. ``Transformation`` merges them into
u = T u_bar on the host (one constraint per slave dof, prescribed slaves dropped, chains
resolved, residual cycles broken) and hands T to the device operator (``exfem_op_set_mpc``),
which then applies opA as free * T^T K T(free * v) + (1 - free) * v in every product with the
boundary mask. Tied contact (surface projection) is not on the device path.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .model import as_dev, stream_ptr, torch, vptr


class MPC:
    """Raw constraint rows: slave dofs (nc,), master dofs (nc, k), weights (nc, k)
    (gfem/constraints.py:14-25)."""

    def __init__(self, slave, masters, weights, name="mpc"):
        self.slave = np.asarray(slave, dtype=np.int64).ravel()
        k = int(np.asarray(masters).size // max(len(self.slave), 1))
        self.masters = np.asarray(masters, dtype=np.int64).reshape(len(self.slave), k)
        self.weights = np.asarray(weights, dtype=np.float64).reshape(len(self.slave), k)
        self.name = name

    def __len__(self):
        return len(self.slave)


def normal_constraint(nodes, normal, name="normal"):
    """n . u = 0 at the nodes (frictionless support or symmetry on any plane,
    gfem/constraints.py:198-209): the dominant normal component is the slave."""
    n = np.asarray(normal, float)
    n = n / np.linalg.norm(n)
    k = int(np.argmax(np.abs(n)))
    others = [j for j in range(3) if j != k]
    nodes = np.asarray(nodes, dtype=np.int64)
    md = np.stack([nodes * 3 + j for j in others], 1)
    w = np.tile([-n[j] / n[k] for j in others], (len(nodes), 1))
    return MPC(nodes * 3 + k, md, w, name)


def coupling(ref_dofs_nodes, nodes, name="coupling"):
    """Kinematic coupling of translations: every node follows node `ref_dofs_nodes`
    (gfem/constraints.py:212-218)."""
    ref = int(ref_dofs_nodes)
    nodes = np.asarray([n for n in nodes if n != ref], dtype=np.int64)
    sd = (nodes[:, None] * 3 + np.arange(3)).ravel()
    md = np.tile(ref * 3 + np.arange(3), len(nodes))[:, None]
    return MPC(sd, md, np.ones((len(sd), 1)), name)


class Transformation:
    """u = T u_bar, sparse, chains resolved (gfem/constraints.py:221-289). The host matrix is
    built as the reference builds it; products with T and T^T run on the device operator it is
    bound to (``bind``)."""

    def __init__(self, ndof, mpcs, fixed=None):
        self.ndof = ndof
        rows = [m for m in mpcs if len(m)]
        self.active = bool(rows)
        self.slave_mask = np.zeros(ndof, bool)
        self.dropped = 0
        self.T = None
        self._op = None
        if not self.active:
            return
        sl = np.concatenate([m.slave for m in rows])
        kmax = max(m.masters.shape[1] for m in rows)
        md = np.concatenate([np.pad(m.masters, ((0, 0), (0, kmax - m.masters.shape[1])), constant_values=-1)
                             for m in rows])
        wt = np.concatenate([np.pad(m.weights, ((0, 0), (0, kmax - m.weights.shape[1]))) for m in rows])
        _, first = np.unique(sl, return_index=True)          # one constraint per slave: first wins
        sl, md, wt = sl[first], md[first], wt[first]
        if fixed is not None:                                # prescribed slaves are dropped
            ok = ~np.asarray(fixed)[sl]
            self.dropped = int((~ok).sum())
            sl, md, wt = sl[ok], md[ok], wt[ok]
        smask = np.zeros(ndof, bool)
        smask[sl] = True
        valid = md >= 0
        keep = np.nonzero(~smask)[0]
        T = sp.csr_matrix((np.concatenate([wt[valid], np.ones(len(keep))]),
                           (np.concatenate([np.repeat(sl, valid.sum(1)), keep]),
                            np.concatenate([md[valid], keep]))), shape=(ndof, ndof))
        for _ in range(8):                                   # resolve chains u_s -> u_s' -> u_m
            if not T[:, smask].nnz:
                break
            T = T @ T
        T = (T @ sp.diags((~smask).astype(float))).tocsr()   # break residual cycles
        T.eliminate_zeros()
        self.T = T
        self.slave_mask = smask

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
