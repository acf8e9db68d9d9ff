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
