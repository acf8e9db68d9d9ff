"""Element types on the device path (gfem/elements.py:113-126,156-199).

The shape-function tables themselves live in the HIP library (csrc/common.h); this module keeps
the host-side descriptors: node count, Gauss-point count, the library's element code and the
Abaqus face tables used to build surface loads.
"""
from . import _lib as L

HEX_FACES = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]]
TET10_FACES = [[0, 1, 2, 4, 5, 6], [0, 3, 1, 7, 8, 4], [1, 3, 2, 8, 9, 5], [2, 3, 0, 9, 7, 6]]


class ElementType:
    def __init__(self, name, nnode, ngp, code, faces, ncorner, fbar=False):
        self.name, self.nnode, self.ngp, self.code = name, nnode, ngp, code
        self.faces = [list(f) for f in faces]
        self.ncorner = ncorner
        self.fbar = fbar

    @property
    def nd(self):
        return 3 * self.nnode

    def __repr__(self):
        return f"<ElementType {self.name}>"


HEX8 = ElementType("C3D8", 8, 8, L.HEX8, HEX_FACES, 8, fbar=True)
HEX8_PLAIN = ElementType("C3D8_FULL", 8, 8, L.HEX8_PLAIN, HEX_FACES, 8)
TET10 = ElementType("C3D10", 10, 4, L.TET10, TET10_FACES, 4)

# gfem/elements.py:194-197, restricted to the element types of the device path
ELEMENTS = {"C3D8": HEX8, "C3D8H": HEX8, "C3D8R": HEX8, "C3D8I": HEX8, "C3D8_FULL": HEX8_PLAIN,
            "C3D10": TET10, "C3D10M": TET10, "C3D10H": TET10}
BY_NNODE = {8: HEX8, 10: TET10}
