"""exfem_amd: the EXFEM (gfem) EbE stiffness operator, PCG and SA-AMG V-cycle on MI355X.

Host code mirrors the reference's Python entry points (gfem.model / gfem.linalg /
gfem.solvers / gfem.optimize); the arithmetic runs in hand-written HIP kernels for gfx950
(libexfem.so, C ABI in include/exfem.h).
"""
from . import _lib
from .elements import ELEMENTS, HEX8, HEX8_PLAIN, TET10
from .materials import (ArrudaBoyce, Gent, J2Plasticity, LinearElastic, MooneyRivlin, NeoHookean,
                        StVenantKirchhoff, Yeoh, lame)
from .model import (Block, EbeOperator, KeBlock, Model, gather_matvec, gather_plan, ops_assemble,
                    ops_diagonal, ops_matvec)
from .solvers import Monitor, State, StaticSolver, Step
from .eigen import buckling, lobpcg, modal
from .constraints import MPC, Transformation, coupling, normal_constraint
from . import mesh

__version__ = "0.1.0"
__all__ = ["Model", "Block", "KeBlock", "EbeOperator", "LinearElastic", "J2Plasticity", "lame", "HEX8",
           "HEX8_PLAIN", "TET10", "ELEMENTS", "StaticSolver", "Step", "MPC", "Transformation", "coupling", "normal_constraint", "State", "Monitor",
           "ops_matvec", "ops_diagonal", "ops_assemble", "gather_plan", "gather_matvec", "mesh",
           "modal", "lobpcg"]
