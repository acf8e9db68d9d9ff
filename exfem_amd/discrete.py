"""Discrete elements: springs, bushings and grounded springs (gfem/discrete.py:1-120).

They follow the element-block interface of the model (dofs, force, stiffness) and enter the
device operator as `rest` operators (exfem_op_set_rest): every product, diagonal and solve on
the operator includes them. Forces and tangents are evaluated on the device
(exfem_spring_eval); nothing here runs arithmetic on the host.

* SpringBlock kind "axial": 2-node spring along the current line of the nodes,
  energy k/2 (L - L0)^2 (large rotation);
* SpringBlock kind "bushing": 2-node spring with stiffnesses (kx, ky, kz) in global axes;
* GroundSpring: node to ground with (kx, ky, kz).
"""
import numpy as np

from . import _lib as L
from .model import as_dev, device, stream_ptr, torch, vptr

_KIND = {"axial": 0, "bushing": 1, "ground": 2}


class _Prop:
    """The 'material' of a discrete block (gfem/discrete.py:22-41): stiffness k, no mass."""
    linear = True
    density = 0.0
    nstate = 0

    def __init__(self, k, name):
        self.k, self.name = k, name
        self.E, self.nu = float(np.max(k)), 0.0

    def describe(self):
        return f"spring(k={self.k})"


class SpringBlock:
    """ne two-node springs (gfem/discrete.py:44-95): kind "axial" (k scalar) or "bushing"
    (k scalar or (kx, ky, kz))."""
    nnode = 2

    def __init__(self, conn, k, kind="axial", name="spring"):
        if kind not in ("axial", "bushing"):
            raise ValueError(f"unknown spring kind {kind!r}")
        self.conn = np.ascontiguousarray(np.asarray(conn, dtype=np.int32).reshape(-1, self.nnode))
        self.ne = len(self.conn)
        self.kind = kind
        kk = np.broadcast_to(np.asarray(k, float), (3,)) if kind == "bushing" else np.full(3, float(k))
        self.k = np.ascontiguousarray(kk, dtype=np.float64)
        self.material = _Prop(k, name)
        self.name = name
        self._conn_dev = None

    def dofs(self):
        return (self.conn.astype(np.int64)[:, :, None] * 3 + np.arange(3)).reshape(self.ne, -1)

    def conn_dev(self):
        if self._conn_dev is None:
            self._conn_dev = torch().as_tensor(self.conn, device=device())
        return self._conn_dev

    def evaluate(self, X_dev, u, scale=None, force=True, stiff=True):
        """(fe (ne, 3 nnode) or None, Ke (ne, 3 nnode, 3 nnode) or None) at u, on the device."""
        nd = 3 * self.nnode
        ud = as_dev(u)
        fe = torch().empty((self.ne, nd), dtype=torch().float64, device=device()) if force else None
        Ke = torch().empty((self.ne, nd, nd), dtype=torch().float64, device=device()) if stiff else None
        sc = None if scale is None else as_dev(scale)
        L.check(L.load().exfem_spring_eval(_KIND[self.kind], self.ne, vptr(X_dev), vptr(self.conn_dev()),
                                           vptr(ud), self.k.ctypes.data, vptr(sc), vptr(fe), vptr(Ke),
                                           stream_ptr()))
        return fe, Ke

    def init_state(self):
        return None


class GroundSpring(SpringBlock):
    """Node-to-ground springs with stiffnesses (kx, ky, kz) (gfem/discrete.py:98-120)."""
    nnode = 1

    def __init__(self, nodes, k, name="ground spring"):
        self.conn = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32).reshape(-1, 1))
        self.ne = len(self.conn)
        self.kind = "ground"
        self.k = np.ascontiguousarray(np.broadcast_to(np.asarray(k, float), (3,)), dtype=np.float64)
        self.material = _Prop(k, name)
        self.name = name
        self._conn_dev = None


__all__ = ["SpringBlock", "GroundSpring"]
