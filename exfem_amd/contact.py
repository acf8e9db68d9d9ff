"""Frictionless penalty contact and rigid walls (gfem/contact.py).

`PenaltyContact`: node-to-surface contact. The candidate master facets of every slave node are
found on the host with a k-d tree (re-searched only after the slaves moved more than a quarter
of the facet size, as gfem/contact.py:126-140 does); the nearest candidate, the gap, the active
flag and the pair's force and Gauss-Newton tangent are evaluated on the device
(exfem_contact_select, exfem_contact_pairs). The pairs enter the linear solve as `rest`
operators (dofs, K) of 12 x 12 (gfem/contact.py:203-212). Friction and augmented-Lagrange
updates are not on the device path.

`RigidWall`: slave nodes against an analytical plane moving with the load factor
(exfem_wall_eval, gfem/contact.py:240-265).
"""
import numpy as np

from . import _lib as L
from .model import as_dev, device, stream_ptr, torch, vptr


class PenaltyContact:
    """Frictionless node-to-surface penalty contact (gfem/contact.py:92-212).
    master_tris: (nt, 3) node ids of the master facets (triangles); slave: node ids (nodes that
    are also master nodes are dropped); stiffness: penalty per slave (scalar or (ns,)); h: facet
    size (depth of captured penetration 0.5 h unless tol)."""
    kind = "contact"

    def __init__(self, X, master_tris, slave, stiffness, h, tol=None, name="contact", ncand=8):
        X = np.asarray(X, float)
        self.tris = np.ascontiguousarray(np.asarray(master_tris, dtype=np.int64).reshape(-1, 3))
        mnodes = set(self.tris.ravel().tolist())
        self.slave = np.array([n for n in np.unique(slave) if n not in mnodes], dtype=np.int64)
        self.ns = len(self.slave)
        self.h = float(h)
        self.depth = float(tol) if tol is not None else 0.5 * self.h
        self.k = np.broadcast_to(np.asarray(stiffness, float), (self.ns,)).copy()
        self.lam = np.zeros(self.ns)
        self.ncand = min(int(ncand), len(self.tris))
        self.name = name
        self._x_search = None
        self._cand = None
        self.searches = 0

    def _search(self, x):
        from scipy.spatial import cKDTree
        cen = x[self.tris].mean(1)
        _, cand = cKDTree(cen).query(x[self.slave], k=self.ncand)
        cand = np.asarray(cand).reshape(self.ns, self.ncand)
        self._cand = torch().as_tensor(self.tris[cand].astype(np.int32), device=device())
        self._x_search = x.copy()
        self.searches += 1

    def params(self, X, u, lam=0.0):
        """Re-search the candidate facets if the slaves moved (host k-d tree), then select the
        pairs on the device: dict(nodes (ns, 4), active (ns,), gap (ns,))."""
        X = np.asarray(X, float)
        uh = u.detach().cpu().numpy() if isinstance(u, torch().Tensor) else np.asarray(u, float)
        x = X + uh.reshape(-1, 3)
        moved = np.inf if self._x_search is None else float(np.abs(x[self.slave] - self._x_search[self.slave]).max())
        if moved > 0.25 * self.h:
            self._search(x)
        t = torch()
        Xd, ud = as_dev(X), as_dev(u)
        sl = t.as_tensor(self.slave.astype(np.int32), device=device())
        nodes = t.empty((self.ns, 4), dtype=t.int32, device=device())
        active = t.empty(self.ns, dtype=t.float64, device=device())
        gap = t.empty(self.ns, dtype=t.float64, device=device())
        L.check(L.load().exfem_contact_select(self.ns, self.ncand, vptr(Xd), vptr(ud), vptr(sl),
                                              vptr(self._cand), self.h, self.depth, vptr(nodes),
                                              vptr(active), vptr(gap), stream_ptr()))
        return dict(nodes=nodes, active=active, gap=gap, X=Xd, u=ud)

    def evaluate(self, p, force=True, stiff=True):
        """(fe (ns, 12), (dofs (ns, 12), K (ns, 12, 12))) of the selected pairs
        (gfem/contact.py:203-212); element force vectors, assembled by the caller."""
        t = torch()
        fe = t.empty((self.ns, 12), dtype=t.float64, device=device()) if force else None
        K = t.empty((self.ns, 12, 12), dtype=t.float64, device=device()) if stiff else None
        kd, ld = as_dev(self.k), as_dev(self.lam)
        L.check(L.load().exfem_contact_pairs(self.ns, vptr(p["X"]), vptr(p["u"]), vptr(p["nodes"]),
                                             vptr(kd), vptr(ld), vptr(p["active"]), vptr(fe), vptr(K),
                                             None, stream_ptr()))
        return fe, K

    def conn(self, p):
        return p["nodes"].cpu().numpy()

    def status(self, p):
        g = p["gap"].cpu().numpy()
        a = p["active"].cpu().numpy()
        pen = -(g * a)[a > 0].min() if (a > 0).any() else 0.0
        return dict(name=self.name, closed=int(((g < 0) & (a > 0)).sum()), max_penetration=float(max(pen, 0)))


class RigidWall:
    """Rigid plane through `point` with outward normal `normal`; slave nodes stay on the positive
    side; the plane moves by `motion` times the load factor (gfem/contact.py:240-265)."""
    kind = "wall"

    def __init__(self, nodes, point, normal, stiffness, motion=(0, 0, 0), name="wall"):
        self.slave = np.asarray(nodes, dtype=np.int64).ravel()
        n = np.asarray(normal, float)
        self.n = n / np.linalg.norm(n)
        self.p0 = np.asarray(point, float)
        self.motion = np.asarray(motion, float)
        self.k = float(stiffness)
        self.name = name
        self.ns = len(self.slave)

    def params(self, X, u, lam=0.0):
        sl = torch().as_tensor(self.slave.astype(np.int32), device=device())
        return dict(nodes=sl, point=np.ascontiguousarray(self.p0 + lam * self.motion, dtype=np.float64),
                    X=as_dev(X), u=as_dev(u))

    def evaluate(self, p, force=True, stiff=True):
        t = torch()
        fe = t.empty((self.ns, 3), dtype=t.float64, device=device()) if force else None
        K = t.empty((self.ns, 3, 3), dtype=t.float64, device=device()) if stiff else None
        nv = np.ascontiguousarray(self.n, dtype=np.float64)
        L.check(L.load().exfem_wall_eval(self.ns, vptr(p["X"]), vptr(p["u"]), vptr(p["nodes"]),
                                         p["point"].ctypes.data, nv.ctypes.data, self.k, vptr(fe),
                                         vptr(K), None, stream_ptr()))
        return fe, K

    def conn(self, p):
        return self.slave.reshape(-1, 1).astype(np.int32)


__all__ = ["PenaltyContact", "RigidWall"]
