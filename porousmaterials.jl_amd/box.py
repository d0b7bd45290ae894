"""Periodic box: the Xtals.jl ``Box`` semantics the energy path relies on.

Xtals.jl is not vendored in the reference; the convention restated here is the
one its doctests print (``docs/src/box.md:28-44``, ``docs/src/crystal.md:48-54``):
``f_to_c`` is upper-triangular with its columns the lattice vectors a, b, c
(a along x), ``c_to_f = inv(f_to_c)`` and ``reciprocal_lattice =
2*pi*inv(f_to_c)`` with the reciprocal vectors as rows
(``src/electrostatics.jl:164,219-220``).
"""
from __future__ import annotations

import math

import numpy as np


class Box:
    """Unit cell with lattice parameters in Angstrom and radians.

    ``Box(a, b, c, alpha, beta, gamma)`` or ``Box(f_to_c)`` (3x3).
    """

    def __init__(self, *args):
        if len(args) == 1:
            f = np.array(args[0], dtype=np.float64).reshape(3, 3)
            self.f_to_c = f
            va, vb, vc = f[:, 0], f[:, 1], f[:, 2]
            self.a = float(np.linalg.norm(va))
            self.b = float(np.linalg.norm(vb))
            self.c = float(np.linalg.norm(vc))
            self.alpha = math.acos(float(np.dot(vb, vc)) / (self.b * self.c))
            self.beta = math.acos(float(np.dot(va, vc)) / (self.a * self.c))
            self.gamma = math.acos(float(np.dot(va, vb)) / (self.a * self.b))
            self.Omega = abs(float(np.linalg.det(f)))
        elif len(args) == 6:
            a, b, c, al, be, ga = (float(x) for x in args)
            self.a, self.b, self.c = a, b, c
            self.alpha, self.beta, self.gamma = al, be, ga
            ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
            omega = a * b * c * math.sqrt(1.0 - ca * ca - cb * cb - cg * cg + 2.0 * ca * cb * cg)
            self.Omega = omega
            self.f_to_c = np.array(
                [[a, b * cg, c * cb],
                 [0.0, b * sg, c * (ca - cb * cg) / sg],
                 [0.0, 0.0, omega / (a * b * sg)]],
                dtype=np.float64,
            )
        else:
            raise TypeError("Box(a, b, c, alpha, beta, gamma) or Box(f_to_c)")
        self.c_to_f = np.linalg.inv(self.f_to_c)
        self.reciprocal_lattice = 2.0 * math.pi * np.linalg.inv(self.f_to_c)

    @property
    def is_upper_triangular(self) -> bool:
        f = self.f_to_c
        return f[1, 0] == 0.0 and f[2, 0] == 0.0 and f[2, 1] == 0.0

    def perpendicular_widths(self) -> np.ndarray:
        """Distance between opposite faces along each lattice direction."""
        f = self.f_to_c
        va, vb, vc = f[:, 0], f[:, 1], f[:, 2]
        n_bc, n_ac, n_ab = np.cross(vb, vc), np.cross(va, vc), np.cross(va, vb)
        return np.array([self.Omega / np.linalg.norm(n_bc),
                         self.Omega / np.linalg.norm(n_ac),
                         self.Omega / np.linalg.norm(n_ab)])

    def __repr__(self) -> str:
        d = 180.0 / math.pi
        return (f"Box(a={self.a:.6f}, b={self.b:.6f}, c={self.c:.6f}, "
                f"alpha={self.alpha * d:.6f} deg, beta={self.beta * d:.6f} deg, "
                f"gamma={self.gamma * d:.6f} deg, Omega={self.Omega:.6f} A^3)")


def unit_cube() -> Box:
    return Box(1.0, 1.0, 1.0, math.pi / 2, math.pi / 2, math.pi / 2)


def replicate_box(box: Box, rep) -> Box:
    """Xtals ``replicate(box, repfactors)``: lengths scaled, angles kept."""
    ra, rb, rc = (int(r) for r in rep)
    return Box(box.a * ra, box.b * rb, box.c * rc, box.alpha, box.beta, box.gamma)
