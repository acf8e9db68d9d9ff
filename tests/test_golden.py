"""The committed golden fixtures agree with the oracle and with the closed forms they encode."""
import os

import numpy as np

from oracle import gfem_oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_c0_fixture_consistent():
    g = np.load(os.path.join(G, "c0_cantilever.npz"))
    X, c = O.box_hex(100, 20, 20, 20, 4, 4)
    K = O.element_stiffness(O.HEX8, X[c[g["Ke_sel"]]], 210000.0, 0.3)
    assert np.allclose(K, g["Ke"], rtol=0, atol=1e-12 * np.abs(K).max())
    assert np.isclose(g["f"].reshape(-1, 3)[:, 2].sum(), -400.0)
    # PCG stopped by the reference rule and agrees with the direct solve
    assert g["pcg_relres"] <= 1e-10
    u = g["u"]
    assert np.linalg.norm(g["pcg_x"] - u) <= 1e-7 * np.linalg.norm(u)
    # Timoshenko cantilever: delta = PL^3/3EI + PL/(kGA), P = 400, L = 100, 20x20 section
    E, nu, P, L_, h = 210000.0, 0.3, 400.0, 100.0, 20.0
    I = h ** 4 / 12
    d_beam = P * L_ ** 3 / (3 * E * I) + P * L_ / (5 / 6 * E / (2 * (1 + nu)) * h * h)
    tip = np.nonzero(X[:, 0] > 100 - 1e-6)[0]
    d = -u.reshape(-1, 3)[tip, 2].mean()
    assert abs(d - d_beam) / d_beam < 0.05


def test_tet10_fixture_consistent():
    g = np.load(os.path.join(G, "tet10_small.npz"))
    K = O.element_stiffness(O.TET10, g["X"][g["conn"][:6]], 1.0, 0.3, g["scale"][:6])
    assert np.allclose(K, g["Ke"], rtol=0, atol=1e-12 * np.abs(K).max())
