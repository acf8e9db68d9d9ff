"""Linear-elastic material (gfem/materials.py:18-19,56-68)."""


def lame(E, nu):
    """(mu, lambda) as gfem/materials.py:18-19."""
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


class LinearElastic:
    """Small-strain isotropic linear elasticity: sigma = lam tr(eps) I + 2 mu eps."""
    small_strain = True
    linear = True
    nstate = 1

    def __init__(self, E, nu, density=1.0, name=None, alpha=0.0, tref=0.0, k=1.0, cp=1.0):
        self.E, self.nu = float(E), float(nu)
        self.mu, self.lam = lame(self.E, self.nu)
        self.density, self.alpha, self.tref = float(density), float(alpha), float(tref)
        self.k, self.cp = float(k), float(cp)
        self.name = name or "material"

    def modulus(self):
        return float(self.mu * (3 * self.lam + 2 * self.mu) / (self.lam + self.mu))

    def describe(self):
        return f"LinearElastic(E={self.E:g}, nu={self.nu:g})"


class J2Plasticity:
    """Small-strain von Mises plasticity with combined hardening (gfem/materials.py:228-300):
    isotropic from a (yield stress, plastic strain) table like Abaqus *PLASTIC - one row is
    perfect plasticity, the last segment extends to eps_p = 1e6 - or Johnson-Cook A + B eps_p^n
    (at T = 0: no thermal softening); kinematic (Prager) with modulus Hk. Radial return with a
    25-step local Newton solve, on the device (exfem_op_set_j2). State per Gauss point:
    [eps_p (9), alpha, backstress (9)]."""
    small_strain = True
    linear = False
    nstate = 19

    def __init__(self, E, nu, table=None, Hk=0.0, johnson_cook=None, density=1.0, name=None,
                 alpha=0.0, tref=0.0, k=1.0, cp=1.0):
        import numpy as np
        self.E, self.nu = float(E), float(nu)
        self.mu, self.lam = lame(self.E, self.nu)
        self.K = self.lam + 2 * self.mu / 3
        self.Hk = float(Hk)
        self.density, self.alpha, self.tref = float(density), float(alpha), float(tref)
        self.k, self.cp = float(k), float(cp)
        self.name = name or "material"
        self.jc = johnson_cook
        if self.jc is None:
            if table is None:
                raise ValueError("J2Plasticity needs a hardening table or johnson_cook")
            t = np.asarray(table, dtype=np.float64).reshape(-1, 2)
            if t.shape[0] > 1 and not np.all(np.diff(t[:, 1]) > 0):
                raise ValueError("the plastic strains of the table must increase")
            self.table = t
            self.sy0 = float(t[0, 0])
        else:
            if "Tmelt" in self.jc or "Ttrans" in self.jc:
                raise NotImplementedError("Johnson-Cook thermal softening is not on the device path")
            self.sy0 = float(self.jc["A"])

    def eqps(self, s):
        """Equivalent plastic strain of a state array (..., 19)."""
        return s[..., 9]

    def modulus(self):
        return self.E

    def describe(self):
        return f"J2Plasticity(E={self.E:g}, nu={self.nu:g}, sy0={self.sy0:g}, Hk={self.Hk:g})"
