"""Writes tests/golden/*.npz from the CPU oracle (oracle/gfem_oracle.py).

    python tests/golden/make_golden.py

Fixtures (small, seeded):
  c0_cantilever.npz : C3D8 20x4x4 cantilever (1,575 dof): Ke of elements 0..7 and 317..319,
                      x ~ N(0,1) rng(1), y = K x (gather order), bc-masked A x, diag(K),
                      Jacobi-PCG (tol 1e-10) iterations / relres, direct solution u.
  tet10_small.npz   : C3D10 box (Kuhn split 3x2x2, jittered), scale per element: Ke of the first
                      6 elements, x, y = K x.
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import gfem_oracle as O  # noqa: E402


def c0():
    X, c = O.box_hex(100, 20, 20, 20, 4, 4)
    nd = 3 * len(X)
    E, nu = 210000.0, 0.3
    K = O.element_stiffness(O.HEX8, X[c], E, nu)
    ops = [(O.dofs(c), K)]
    x = np.random.default_rng(1).standard_normal(nd)
    y = O.gather_matvec(O.gather_plan([c], len(X)), ops, x, nd)
    free = np.ones(nd)
    fix = np.nonzero(X[:, 0] < 1e-9)[0]
    for d in range(3):
        free[fix * 3 + d] = 0
    yb = O.op_bc(lambda w: O.gather_matvec(O.gather_plan([c], len(X)), ops, w, nd), free)(x)
    diag = O.ops_diagonal(ops, nd)
    tip = X[:, 0] > 100 - 1e-6
    faces = c[:, [1, 2, 6, 5]]
    faces = faces[tip[faces].all(1)]
    f = O.traction_vector(O.quad4_weights(X, faces), (0, 0, -1.0))
    b = free * f
    A = O.op_bc(lambda w: O.gather_matvec(O.gather_plan([c], len(X)), ops, w, nd), free)
    inv = O.jacobi_inv(diag, free)
    xs, k, rel = O.pcg(A, b, lambda r: inv * r, tol=1e-10, maxiter=20000)
    u = spla.spsolve(O.bc_matrix(O.ops_assemble(ops, nd), free).tocsc(), b)
    sel = np.r_[0:8, 317:320]
    np.savez_compressed(os.path.join(HERE, "c0_cantilever.npz"), Ke_sel=sel, Ke=K[sel], x=x, y=y,
                        y_bc=yb, diag=diag, free=free, f=f, pcg_iters=k, pcg_relres=rel,
                        pcg_x=xs, u=u)


def tet10():
    X, c = O.box_hex(3.0, 2.0, 2.0, 3, 2, 2)
    X = X + np.random.default_rng(4).uniform(-0.1, 0.1, X.shape) * \
        ((X > 1e-9) & (X < np.array([3, 2, 2]) - 1e-9)).all(1)[:, None]
    X, c = O.hex_to_tet4(X, c)
    X, c = O.tet4_to_tet10(X, c)
    sc = np.exp(np.random.default_rng(5).uniform(np.log(1e3), np.log(1e6), len(c)))
    K = O.element_stiffness(O.TET10, X[c], 1.0, 0.3, sc)
    nd = 3 * len(X)
    x = np.random.default_rng(6).standard_normal(nd)
    y = O.gather_matvec(O.gather_plan([c], len(X)), [(O.dofs(c), K)], x, nd)
    np.savez_compressed(os.path.join(HERE, "tet10_small.npz"), X=X, conn=c, scale=sc, Ke=K[:6],
                        x=x, y=y)


if __name__ == "__main__":
    c0()
    tet10()
    print("written", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
