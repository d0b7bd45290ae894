"""Rigid adsorbate (probe) model.

Mirrors ``Molecule{T}`` (``src/molecule.jl:10-15``): LJ spheres (``atoms``)
and point charges (``charges``) with their own coordinates, a centre of mass,
Cartesian or fractional coordinates (``Frac``/``Cart``, ``:100-109``),
``translate_to!`` (``:171-183``), ``ion`` (``:340-343``), the Arvo random
rotation (``:233-246``) and ``needs_rotations`` (``src/mc_helpers.jl:8``).
"""
from __future__ import annotations

import copy
import csv
import math
import os

import numpy as np

from .box import Box
from .crystal import ATOMIC_MASSES


class Molecule:
    """``atoms_x`` (3, n_lj) and ``charges_x`` (3, n_q) hold Cartesian
    coordinates when ``frac`` is False, fractional ones otherwise."""

    def __init__(self, species, atom_species, atoms_x, charges_q, charges_x, com, frac=False):
        self.species = str(species)
        self.atom_species = [str(s) for s in atom_species]
        self.atoms_x = np.array(atoms_x, dtype=np.float64).reshape(3, -1)
        self.charges_q = np.array(charges_q, dtype=np.float64).reshape(-1)
        self.charges_x = np.array(charges_x, dtype=np.float64).reshape(3, -1)
        self.com = np.array(com, dtype=np.float64).reshape(3)
        self.frac = bool(frac)

    @classmethod
    def from_dir(cls, path: str, check_neutrality: bool = True) -> "Molecule":
        """Read ``atoms.csv`` (atom,x,y,z) and ``charges.csv`` (q,x,y,z)
        (``src/molecule.jl:56-91``)."""
        def rows(fn):
            with open(os.path.join(path, fn)) as f:
                return list(csv.DictReader([ln for ln in f if ln.strip()]))
        a = rows("atoms.csv")
        c = rows("charges.csv")
        species = [r["atom"].strip() for r in a]
        ax = np.array([[float(r[k]) for r in a] for k in "xyz"]).reshape(3, -1)
        q = np.array([float(r["q"]) for r in c])
        cx = np.array([[float(r[k]) for r in c] for k in "xyz"]).reshape(3, -1)
        m = cls(os.path.basename(os.path.normpath(path)), species, ax, q, cx, np.zeros(3))
        m.com = center_of_mass(m)
        if check_neutrality and abs(float(np.sum(q))) > 0.001:
            raise ValueError(f"Molecule {m.species} is not charge neutral!")
        return m

    @property
    def n_atoms(self) -> int:
        return self.atoms_x.shape[1]

    @property
    def n_charges(self) -> int:
        return self.charges_q.size

    def copy(self) -> "Molecule":
        return copy.deepcopy(self)

    def __repr__(self) -> str:
        return (f"Molecule({self.species!r}, {self.n_atoms} LJ spheres, {self.n_charges} charges, "
                f"{'Frac' if self.frac else 'Cart'})")


def center_of_mass(m: Molecule) -> np.ndarray:
    """``src/molecule.jl:24-39`` (unweighted mean if all masses are zero)."""
    masses = np.array([ATOMIC_MASSES.get(s, 0.0) for s in m.atom_species])
    if m.n_atoms == 0:
        return np.mean(m.charges_x, axis=1) if m.n_charges else np.zeros(3)
    if masses.sum() == 0.0:
        return np.mean(m.atoms_x, axis=1)
    return (m.atoms_x * masses[None, :]).sum(axis=1) / masses.sum()


def needs_rotations(m: Molecule) -> bool:
    return m.n_atoms + m.n_charges > 1


def has_charges(m: Molecule) -> bool:
    return m.n_charges > 0


def Frac(m: Molecule, box: Box) -> Molecule:
    if m.frac:
        return m.copy()
    out = m.copy()
    out.atoms_x = box.c_to_f @ m.atoms_x
    out.charges_x = box.c_to_f @ m.charges_x
    out.com = box.c_to_f @ m.com
    out.frac = True
    return out


def Cart(m: Molecule, box: Box) -> Molecule:
    if not m.frac:
        return m.copy()
    out = m.copy()
    out.atoms_x = box.f_to_c @ m.atoms_x
    out.charges_x = box.f_to_c @ m.charges_x
    out.com = box.f_to_c @ m.com
    out.frac = False
    return out


def translate_to(m: Molecule, x) -> None:
    """Move the centre of mass to ``x`` (same coordinate kind as ``m``)."""
    d = np.asarray(x, dtype=np.float64).reshape(3) - m.com
    m.atoms_x = m.atoms_x + d[:, None]
    m.charges_x = m.charges_x + d[:, None]
    m.com = m.com + d


def ion(q: float, xf) -> Molecule:
    """Point-charge probe at fractional ``xf`` (``src/molecule.jl:340-343``)."""
    xf = np.asarray(xf, dtype=np.float64).reshape(3)
    return Molecule("ion", [], np.zeros((3, 0)), [float(q)], xf.reshape(3, 1), xf, frac=True)


def arvo_rotation_matrix(u1: float, u2: float, u3: float) -> np.ndarray:
    """Arvo's random rotation from three uniforms in [0, 1)
    (``src/molecule.jl:233-244``): -H * Rz(2 pi u1), H = I - 2 v v^T."""
    a = u1 * 2.0 * math.pi
    r = np.array([[math.cos(a), math.sin(a), 0.0], [-math.sin(a), math.cos(a), 0.0], [0.0, 0.0, 1.0]])
    b = 2.0 * math.pi * u2
    v = np.array([math.cos(b) * math.sqrt(u3), math.sin(b) * math.sqrt(u3), math.sqrt(1.0 - u3)])
    h = np.eye(3) - 2.0 * np.outer(v, v)
    return -h @ r


def site_offsets(m: Molecule):
    """Cartesian offsets of LJ spheres and charges from the centre of mass
    (the rigid template the device rotates per insertion)."""
    if m.frac:
        raise ValueError("site_offsets needs Cartesian coordinates")
    return m.atoms_x - m.com[:, None], m.charges_x - m.com[:, None]


def point_charge(q: float = 1.0, species: str = "ion") -> Molecule:
    """Cartesian single point-charge probe centred at the origin (the
    template form of ``ion``, usable with ``energy_grid``)."""
    return Molecule(species, [], np.zeros((3, 0)), [float(q)], np.zeros((3, 1)), np.zeros(3))


def lj_sphere(species: str) -> Molecule:
    """Cartesian single LJ-sphere probe (e.g. methane ``CH4``, ``Xe``)."""
    return Molecule(species, [species], np.zeros((3, 1)), [], np.zeros((3, 0)), np.zeros(3))
