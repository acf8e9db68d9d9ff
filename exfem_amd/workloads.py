"""The benchmark workloads of the solver path, built on the host.

* cantilever C3D8 box 20x4x4 of 100x20x20, clamped at x=0, tip traction (0,0,-1) on x=100
  (525 nodes, 1,575 dof);
* the same box at 200x40x40 (337,881 nodes, 1,013,643 dof) with interior nodes jittered by
  U(-10%, +10%) of the element size per component (numpy default_rng(seed));
* a C3D10 box 1x1x4 from a seeded Delaunay of a jittered lattice (boundary nodes jittered
  tangentially only, so no hull tet is flat), E log-uniform in [1e3, 1e6] per element;
* 64 load cases of Gaussian nodal forces on the top face (z = W) of the jittered box;
* a stream of solve requests on C3D8 boxes with log-uniform dof counts.
"""
import numpy as np

from . import mesh as M
from .elements import HEX8, TET10
from .loads import face_facets, traction_vector


class Problem:
    """Nodes, one element block, material, boundary mask and load(s) of a workload."""

    def __init__(self, name, X, conn, etype, E, nu, free, f, scale=None, meta=None):
        self.name, self.X, self.conn, self.etype = name, X, conn, etype
        self.E, self.nu, self.free, self.f, self.scale = E, nu, free, f, scale
        self.meta = meta or {}

    @property
    def nn(self):
        return len(self.X)

    @property
    def ndof(self):
        return 3 * len(self.X)

    @property
    def ne(self):
        return len(self.conn)

    def model(self):
        from .materials import LinearElastic
        from .model import Model
        m = Model(self.X)
        m.add_block(self.conn, self.etype, LinearElastic(self.E, self.nu))
        return m


def _clamp_x0(X):
    free = np.ones(3 * len(X), np.uint8)
    fix = M.nodes_where(X, lambda x, y, z: x < 1e-9)
    for d in range(3):
        free[fix * 3 + d] = 0
    return free


def cantilever(nx=20, ny=4, nz=4, L=100.0, H=20.0, W=20.0, E=210000.0, nu=0.3, jitter=0.0,
               seed=0, traction=(0.0, 0.0, -1.0), etype=HEX8):
    X, conn = M.box_hex(L, H, W, nx, ny, nz)
    if jitter:
        h = np.array([L / nx, H / ny, W / nz])
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, X.shape) * h[None, :]
        eps = 1e-9 * max(L, H, W)
        interior = ((X[:, 0] > eps) & (X[:, 0] < L - eps) & (X[:, 1] > eps) & (X[:, 1] < H - eps)
                    & (X[:, 2] > eps) & (X[:, 2] < W - eps))
        X = X + d * interior[:, None]
    tip = np.zeros(len(X), bool)
    tip[M.nodes_where(X, lambda x, y, z: x > L - 1e-9 * L)] = True
    f = traction_vector(X, face_facets(conn, etype, tip), traction)
    return Problem(f"C3D8 {nx}x{ny}x{nz}", X, conn.astype(np.int32), etype, E, nu, _clamp_x0(X), f,
                   meta=dict(L=L, H=H, W=W, nx=nx, ny=ny, nz=nz, jitter=jitter, seed=seed))


def small_cantilever():
    """C3D8 20x4x4 of 100x20x20, 1,575 dof, plain (no jitter)."""
    return cantilever(20, 4, 4)


def jittered_cantilever(nx=200, ny=40, nz=40, seed=0):
    """C3D8 200x40x40 of 100x20x20 with interior jitter +-10% h: 1,013,643 dof."""
    return cantilever(nx, ny, nz, jitter=0.1, seed=seed)


def morton_order(P, bits=10):
    """Permutation sorting points P (n,3) along a 3-D Morton (Z-order) curve."""
    lo, hi = P.min(0), P.max(0)
    q = ((P - lo) / np.maximum(hi - lo, 1e-300) * (2 ** bits - 1)).astype(np.uint64)
    code = np.zeros(len(P), np.uint64)
    for b in range(bits):
        for d in range(3):
            code |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
    return np.argsort(code, kind="stable")


def tet10_box(nx=42, ny=42, nz=168, seed=0, E_range=(1e3, 1e6), nu=0.3, size=(1.0, 1.0, 4.0)):
    """Seeded Delaunay C3D10 box; boundary nodes jittered tangentially only; per-element
    stiffness scale E_e = exp(U(ln 1e3, ln 1e6)) on a base E = 1. Clamped at z = 0,
    traction (-1, 0, 0) on z = size[2]."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = size
    xs, ys, zs = (np.linspace(0, Lx, nx + 1), np.linspace(0, Ly, ny + 1), np.linspace(0, Lz, nz + 1))
    X = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"), -1).reshape(-1, 3)
    h = np.array([Lx / nx, Ly / ny, Lz / nz])
    d = rng.uniform(-0.1, 0.1, X.shape) * h[None, :]
    eps = 1e-9
    on = np.stack([(X[:, 0] < eps) | (X[:, 0] > Lx - eps), (X[:, 1] < eps) | (X[:, 1] > Ly - eps),
                   (X[:, 2] < eps) | (X[:, 2] > Lz - eps)], 1)
    d[on] = 0.0                               # no motion normal to a face the node lies on
    X = X + d
    tets = Delaunay(X).simplices.astype(np.int64)
    tets = M.orient_tets(X, tets)
    tets = tets[morton_order(X[tets].mean(1))]      # spatially coherent element order
    X10, conn = M.tet4_to_tet10(X, tets)
    ne = len(conn)
    scale = np.exp(rng.uniform(np.log(E_range[0]), np.log(E_range[1]), ne))
    free = np.ones(3 * len(X10), np.uint8)
    fix = M.nodes_where(X10, lambda x, y, z: z < 1e-9)
    for k in range(3):
        free[fix * 3 + k] = 0
    top = np.zeros(len(X10), bool)
    top[M.nodes_where(X10, lambda x, y, z: z > Lz - 1e-9)] = True
    f = traction_vector(X10, face_facets(conn, TET10, top), (-1.0, 0.0, 0.0))
    return Problem(f"C3D10 Delaunay {nx}x{ny}x{nz}", X10, conn.astype(np.int32), TET10, 1.0, nu, free,
                   f, scale=scale, meta=dict(seed=seed, lattice=(nx, ny, nz)))


def top_face_loads(p, ncase=64):
    """ncase load cases: N(0,1) nodal forces on every top-face node (z = W), rng(k) for case k."""
    W = p.meta.get("W", p.X[:, 2].max())
    top = M.nodes_where(p.X, lambda x, y, z: z > W - 1e-9 * W)
    F = np.zeros((ncase, p.ndof))
    for k in range(ncase):
        F[k].reshape(-1, 3)[top] = np.random.default_rng(k).standard_normal((len(top), 3))
    return F


def serving_requests(n=512, seed=0, dof_range=(1e4, 2e6), rate=None):
    """Request stream: target dof log-uniform, aspect a ~ U(1, 10) of a W=H=1, L=a box with
    n elements across, E ~ U(70000, 210000), nu 0.3, clamped at x=0, unit tip traction in a random
    direction on x=L; Poisson arrival times at `rate` requests/s (None: all at t=0)."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(n):
        target = np.exp(rng.uniform(np.log(dof_range[0]), np.log(dof_range[1])))
        a = rng.uniform(1.0, 10.0)
        # 3 (a n + 1)(n + 1)^2 ~ target
        nc = max(1, int(round((target / (3 * a)) ** (1 / 3))))
        nx = max(1, int(round(a * nc)))
        E = rng.uniform(70000.0, 210000.0)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        if rate:
            t += rng.exponential(1.0 / rate)
        reqs.append(dict(id=i, nx=nx, ny=nc, nz=nc, L=float(a), E=float(E), direction=d, arrival=t))
    return reqs


def request_problem(r):
    return cantilever(r["nx"], r["ny"], r["nz"], L=r["L"], H=1.0, W=1.0, E=r["E"],
                      traction=tuple(r["direction"]))
