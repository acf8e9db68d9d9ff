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


class _Hyperelastic:
    """Strain-energy material (gfem/materials.py:124-141) on the device (exfem_op_set_hyper):
    P = dW/dF and its tangent from dual / hyper-dual evaluations of W, finite strain (F-bar for
    C3D8). Stateless. mu, lam are the small-strain constants (contact penalty scaling, the
    linear limit at F = I)."""
    small_strain = False
    linear = False
    nstate = 1
    kind = -1

    def _common(self, density, name, alpha, tref, k, cp):
        self.density, self.alpha, self.tref = float(density), float(alpha), float(tref)
        self.k, self.cp = float(k), float(cp)
        self.name = name or "material"
        # small-strain E, nu of (mu, lam): the operator's geometry record needs them
        self.E = float(self.mu * (3 * self.lam + 2 * self.mu) / (self.lam + self.mu))
        self.nu = float(self.lam / (2 * (self.lam + self.mu)))

    def params(self):
        raise NotImplementedError

    def modulus(self):
        return self.E

    def describe(self):
        return f"{type(self).__name__}({', '.join(f'{v:g}' for v in self.params())})"


class NeoHookean(_Hyperelastic):
    """Abaqus form W = C10 (I1bar - 3) + (J - 1)^2 / D1 (gfem/materials.py:146-160); from (E, nu)
    C10 = mu/2, D1 = 2/(lam + 2 mu/3)."""
    kind = 0

    def __init__(self, C10=None, D1=None, E=None, nu=None, density=1.0, name=None, alpha=0.0,
                 tref=0.0, k=1.0, cp=1.0):
        if C10 is None:
            mu, lam = lame(E, nu)
            C10, D1 = mu / 2, 2 / (lam + 2 * mu / 3)
        self.C10, self.D1 = float(C10), float(D1)
        self.mu = 2 * self.C10
        self.lam = 2 / self.D1 - 2 * self.mu / 3
        self._common(density, name, alpha, tref, k, cp)

    def params(self):
        return (self.C10, self.D1)


class MooneyRivlin(_Hyperelastic):
    """W = C10 (I1b - 3) + C01 (I2b - 3) + (J - 1)^2 / D1 (gfem/materials.py:163-174)."""
    kind = 1

    def __init__(self, C10, C01, D1, density=1.0, name=None, alpha=0.0, tref=0.0, k=1.0, cp=1.0):
        self.C10, self.C01, self.D1 = float(C10), float(C01), float(D1)
        self.mu = 2 * (self.C10 + self.C01)
        self.lam = 2 / self.D1 - 2 * self.mu / 3
        self._common(density, name, alpha, tref, k, cp)

    def params(self):
        return (self.C10, self.C01, self.D1)


class Yeoh(_Hyperelastic):
    """W = sum Ci0 (I1b - 3)^i + sum (J - 1)^2i / Di, Di > 0 (gfem/materials.py:177-191)."""
    kind = 2

    def __init__(self, C10, C20, C30, D1, D2=0.0, D3=0.0, density=1.0, name=None, alpha=0.0,
                 tref=0.0, k=1.0, cp=1.0):
        self.C = [float(C10), float(C20), float(C30)]
        self.D = [float(D1), float(D2), float(D3)]
        self.mu = 2 * self.C[0]
        self.lam = 2 / self.D[0] - 2 * self.mu / 3
        self._common(density, name, alpha, tref, k, cp)

    def params(self):
        return tuple(self.C + self.D)


class ArrudaBoyce(_Hyperelastic):
    """Eight-chain model: mu, locking stretch lam_m, D (gfem/materials.py:194-207)."""
    kind = 3

    def __init__(self, mu, lam_m, D, density=1.0, name=None, alpha=0.0, tref=0.0, k=1.0, cp=1.0):
        self.mu0, self.lm, self.D = float(mu), float(lam_m), float(D)
        self.mu = self.mu0
        self.lam = 2 / self.D - 2 * self.mu / 3
        self._common(density, name, alpha, tref, k, cp)

    def params(self):
        return (self.mu0, self.lm, self.D)


class Gent(_Hyperelastic):
    """W = -mu Jm/2 ln(1 - (I1b - 3)/Jm) + ((J^2 - 1)/2 - ln J)/D (gfem/materials.py:210-225)."""
    kind = 4

    def __init__(self, mu, Jm, D, density=1.0, name=None, alpha=0.0, tref=0.0, k=1.0, cp=1.0):
        self.mu, self.Jm, self.D = float(mu), float(Jm), float(D)
        self.lam = 2 / self.D - 2 * self.mu / 3
        self._common(density, name, alpha, tref, k, cp)

    def params(self):
        return (self.mu, self.Jm, self.D)


class StVenantKirchhoff(_Hyperelastic):
    """Large-rotation linear elasticity: P = F S(E), E Green-Lagrange (gfem/materials.py:71-80),
    on the device as the energy lam/2 tr(E)^2 + mu E:E."""
    kind = 5

    def __init__(self, E, nu, density=1.0, name=None, alpha=0.0, tref=0.0, k=1.0, cp=1.0):
        self.mu, self.lam = lame(float(E), float(nu))
        self._common(density, name, alpha, tref, k, cp)

    def params(self):
        return (self.mu, self.lam)
