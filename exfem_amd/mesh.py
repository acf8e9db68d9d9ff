"""Structured mesh generators and converters (gfem/mesh.py:5-42), host side."""
import numpy as np

_EDGES10 = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
_KUHN = np.array([[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]])


def box_hex(L, H, W, nx, ny, nz, origin=(0, 0, 0)):
    """Hex box with node id (i*(ny+1)+j)*(nz+1)+k and Abaqus C3D8 node order."""
    xs = np.linspace(0, L, nx + 1) + origin[0]
    ys = np.linspace(0, H, ny + 1) + origin[1]
    zs = np.linspace(0, W, nz + 1) + origin[2]
    X = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"), -1).reshape(-1, 3)
    i, j, k = (a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                                              indexing="ij"))
    s1, s0 = nz + 1, (ny + 1) * (nz + 1)
    base = i * s0 + j * s1 + k
    off = np.array([0, s0, s0 + s1, s1, 1, s0 + 1, s0 + s1 + 1, s1 + 1])
    return X, base[:, None] + off[None, :]


def hex_to_tet4(X, conn):
    """Kuhn split into 6 tets per hex, flipped to positive volume."""
    tets = np.asarray(conn)[:, _KUHN].reshape(-1, 4)
    return X, orient_tets(X, tets)


def orient_tets(X, tets):
    v = np.einsum("ij,ij->i", np.cross(X[tets[:, 1]] - X[tets[:, 0]], X[tets[:, 2]] - X[tets[:, 0]]),
                  X[tets[:, 3]] - X[tets[:, 0]])
    tets = tets.copy()
    neg = v < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return tets


def tet4_to_tet10(X, conn):
    """One shared mid-edge node per edge, Abaqus C3D10 order (edges 01,12,20,03,13,23)."""
    conn = np.asarray(conn)
    e = np.concatenate([np.sort(conn[:, list(p)], 1) for p in _EDGES10])
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    mids = 0.5 * (X[uniq[:, 0]] + X[uniq[:, 1]])
    newid = inv.reshape(len(_EDGES10), -1).T + len(X)
    return np.vstack([X, mids]), np.hstack([conn, newid])


def nodes_where(X, fn):
    return np.nonzero(fn(X[:, 0], X[:, 1], X[:, 2]))[0]
