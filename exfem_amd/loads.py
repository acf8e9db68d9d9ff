"""Dead surface loads on the host: nodal weights w_a = int N_a dA of boundary facets and the
traction vector (Surface.integrate, gfem/surfaces.py:103-120; traction_vector,
gfem/loads.py:54-57; cload_vector, gfem/loads.py:33-36)."""
import numpy as np

_G = 1 / np.sqrt(3)
_Q4_PTS = [(-_G, -_G), (_G, -_G), (_G, _G), (-_G, _G)]
_QC = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float)
_T3_PTS = [(1 / 6, 1 / 6), (2 / 3, 1 / 6), (1 / 6, 2 / 3)]


def face_facets(conn, etype, node_mask):
    """Element faces (node lists, etype.faces order) whose nodes all satisfy node_mask."""
    conn = np.asarray(conn)
    out = []
    for f in etype.faces:
        fn = conn[:, f]
        sel = node_mask[fn].all(1)
        out.append(fn[sel])
    nf = max(len(f) for f in etype.faces)
    return np.concatenate(out).reshape(-1, nf) if out else np.zeros((0, nf), np.int64)


def _quad4(xe):
    w = np.zeros(xe.shape[:2])
    for s, t in _Q4_PTS:
        N = 0.25 * (1 + _QC[:, 0] * s) * (1 + _QC[:, 1] * t)
        a1 = np.einsum("fai,a->fi", xe, 0.25 * _QC[:, 0] * (1 + _QC[:, 1] * t))
        a2 = np.einsum("fai,a->fi", xe, 0.25 * _QC[:, 1] * (1 + _QC[:, 0] * s))
        w += np.linalg.norm(np.cross(a1, a2), axis=1)[:, None] * N[None, :]
    return w


def _tri6(xe):
    w = np.zeros(xe.shape[:2])
    for s, t in _T3_PTS:
        L = np.array([1 - s - t, s, t])
        dL = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
        N = np.concatenate([L * (2 * L - 1), [4 * L[0] * L[1], 4 * L[1] * L[2], 4 * L[2] * L[0]]])
        dN = np.zeros((6, 2))
        for a in range(3):
            dN[a] = (4 * L[a] - 1) * dL[a]
        for k, (i, j) in enumerate([(0, 1), (1, 2), (2, 0)]):
            dN[3 + k] = 4 * (L[j] * dL[i] + L[i] * dL[j])
        a1 = np.einsum("fai,a->fi", xe, dN[:, 0])
        a2 = np.einsum("fai,a->fi", xe, dN[:, 1])
        w += (np.linalg.norm(np.cross(a1, a2), axis=1) / 6.0)[:, None] * N[None, :]
    return w


def facet_weights(X, facets):
    """Nodal weights int N_a dA for QUAD4 (2x2 Gauss) or TRI6 (3-point rule) facets."""
    X = np.asarray(X, float)
    facets = np.asarray(facets)
    w = np.zeros(len(X))
    if len(facets) == 0:
        return w
    xe = X[facets]
    if facets.shape[1] == 4:
        wf = _quad4(xe)
    elif facets.shape[1] == 6:
        wf = _tri6(xe)
    else:
        raise ValueError(f"unsupported facet with {facets.shape[1]} nodes")
    np.add.at(w, facets, wf)
    return w


def traction_vector(X, facets, t):
    """Uniform traction t (force/area) on the facets, node-interleaved (ndof,)."""
    w = facet_weights(X, facets)
    return (w[:, None] * np.asarray(t, float)[None]).ravel()


def cload_vector(ndof, nodes, dof, value):
    f = np.zeros(ndof)
    np.add.at(f, np.asarray(nodes, dtype=np.int64) * 3 + dof, value)
    return f
