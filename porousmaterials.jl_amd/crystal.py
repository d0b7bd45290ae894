"""Crystal: P1 framework atoms and point charges in fractional coordinates.

Restates the parts of Xtals.jl's ``Crystal`` the energy path uses (Xtals is an
external, un-vendored dependency of the reference, ``Project.toml:28,46``):

* CIF reading of P1 structures: cell lengths/angles (degrees -> radians), the
  ``_atom_site`` loop, species from ``_atom_site_type_symbol`` falling back to
  ``_atom_site_label``, charges from ``_atom_site_charge``, fractional
  coordinates wrapped into [0, 1), zero charges left out of ``charges``;
* ``replicate(crystal, repfactors)`` with the loop order ra, rb, rc, atoms and
  ``xf_new = (xf + [ra, rb, rc]) ./ repfactors`` (``src/henry.jl:86``,
  ``src/grid.jl:332``);
* ``strip_numbers_from_atom_labels!`` (``docs/src/crystal.md:30-40``);
* ``crystal_density`` (``src/henry.jl:96``) from a mass table.
"""
from __future__ import annotations

import math
import os
import re
import shlex

import numpy as np

from .box import Box, replicate_box

# Standard atomic weights (IUPAC 2013 abridged values) for crystal densities;
# extend with ``register_atomic_masses``.
ATOMIC_MASSES = {
    "H": 1.00794, "He": 4.002602, "Li": 6.941, "Be": 9.012182, "B": 10.811, "C": 12.0107,
    "N": 14.0067, "O": 15.9994, "F": 18.9984032, "Ne": 20.1797, "Na": 22.98976928,
    "Mg": 24.305, "Al": 26.9815386, "Si": 28.0855, "P": 30.973762, "S": 32.065,
    "Cl": 35.453, "Ar": 39.948, "K": 39.0983, "Ca": 40.078, "Sc": 44.955912, "Ti": 47.867,
    "V": 50.9415, "Cr": 51.9961, "Mn": 54.938045, "Fe": 55.845, "Co": 58.933195,
    "Ni": 58.6934, "Cu": 63.546, "Zn": 65.38, "Ga": 69.723, "Ge": 72.64, "As": 74.9216,
    "Se": 78.96, "Br": 79.904, "Kr": 83.798, "Rb": 85.4678, "Sr": 87.62, "Y": 88.90585,
    "Zr": 91.224, "Nb": 92.90638, "Mo": 95.96, "Ru": 101.07, "Rh": 102.9055, "Pd": 106.42,
    "Ag": 107.8682, "Cd": 112.411, "In": 114.818, "Sn": 118.71, "Sb": 121.76, "Te": 127.6,
    "I": 126.90447, "Xe": 131.293, "Cs": 132.9054519, "Ba": 137.327, "La": 138.90547,
    "Ce": 140.116, "Nd": 144.242, "Eu": 151.964, "Gd": 157.25, "Tb": 158.92535,
    "Dy": 162.5, "Er": 167.259, "Yb": 173.054, "Hf": 178.49, "W": 183.84, "Pt": 195.084,
    "Au": 196.966569, "Hg": 200.59, "Pb": 207.2, "Bi": 208.9804, "U": 238.02891,
    # pseudo-atoms of the reference's adsorbate models (src/atomic_masses.jl:2-25)
    "N_in_N2": 14.0067, "CH2": 14.025, "CH3": 15.035, "CH4": 16.04, "C_b": 12.0107,
    "C_tol": 12.0107, "C_ac": 12.0107, "C_RCOO": 12.0107, "C_sp2": 12.0107,
    "C_sp3": 12.0107, "C_CO2": 12.0107, "H_H2S": 1.00794, "H_b": 1.00794,
    "O_RCOO": 15.9994, "O_CO2": 15.9994, "O_zeo": 15.9994, "S_H2S": 32.065,
    "Si_zeo": 28.0855, "ig": 1.0, "X": 1.0,
}


def register_atomic_masses(masses: dict) -> None:
    ATOMIC_MASSES.update({str(k): float(v) for k, v in masses.items()})


class Crystal:
    """Framework: ``species`` (list of str) and ``xf`` (3, N) fractional
    coordinates of the atoms, and ``atom_q`` (N,) the point charge each atom
    carries (0 for none).  ``charges_q`` / ``charges_xf`` list the nonzero
    charges, as Xtals keeps them (``include_zero_charges=false``)."""

    def __init__(self, name, box: Box, species, xf, atom_q=None, symmetry_p1: bool = True):
        self.name = str(name)
        self.box = box
        self.species = [str(s) for s in species]
        self.xf = np.array(xf, dtype=np.float64).reshape(3, -1)
        if len(self.species) != self.xf.shape[1]:
            raise ValueError("species and coordinates differ in length")
        self.atom_q = (np.zeros(self.xf.shape[1]) if atom_q is None
                       else np.array(atom_q, dtype=np.float64).reshape(-1))
        if self.atom_q.size != self.xf.shape[1]:
            raise ValueError("atom_q and coordinates differ in length")
        self.symmetry_p1 = symmetry_p1
        # provenance of a replica: (unit-cell crystal, repfactors)
        self._unit = None
        self._rep = (1, 1, 1)
        # per-crystal cache of device handles (see device.py)
        self._device_cache = {}

    @property
    def n_atoms(self) -> int:
        return self.xf.shape[1]

    @property
    def charges_q(self) -> np.ndarray:
        return self.atom_q[self.atom_q != 0.0]

    @property
    def charges_xf(self) -> np.ndarray:
        return self.xf[:, self.atom_q != 0.0]

    def __repr__(self) -> str:
        return f"Crystal({self.name!r}, {self.n_atoms} atoms, {self.charges_q.size} charges, {self.box!r})"


def has_charges(crystal: Crystal) -> bool:
    return crystal.charges_q.size > 0


def assert_P1_symmetry(crystal: Crystal) -> None:
    if not crystal.symmetry_p1:
        raise ValueError(f"crystal {crystal.name} is not in P1 symmetry")


def net_charge(crystal: Crystal) -> float:
    return float(np.sum(crystal.charges_q))


def crystal_density(crystal: Crystal) -> float:
    """Mass density in kg/m^3 (Xtals ``crystal_density``)."""
    try:
        mass = sum(ATOMIC_MASSES[s] for s in crystal.species)  # g/mol
    except KeyError as e:
        raise KeyError(f"no atomic mass for species {e.args[0]}; use register_atomic_masses") from None
    return mass / crystal.box.Omega * 1660.53906660  # amu/A^3 -> kg/m^3


def replicate(crystal: Crystal, repfactors) -> Crystal:
    ra_n, rb_n, rc_n = (int(r) for r in repfactors)
    shifts = np.array([[ra, rb, rc] for ra in range(ra_n) for rb in range(rb_n) for rc in range(rc_n)],
                      dtype=np.float64)
    scale = np.array([ra_n, rb_n, rc_n], dtype=np.float64)[:, None]

    def rep(xf):
        if xf.shape[1] == 0:
            return np.zeros((3, 0))
        blocks = [(xf + s[:, None]) / scale for s in shifts]
        return np.concatenate(blocks, axis=1)

    out = Crystal(crystal.name, replicate_box(crystal.box, repfactors),
                  crystal.species * len(shifts), rep(crystal.xf),
                  np.tile(crystal.atom_q, len(shifts)), crystal.symmetry_p1)
    # the device sees `crystal` as the unit cell and these factors: energy-grid
    # points and accessibility lookups are fractional coordinates of the cell
    # that was replicated, also when that cell is itself a replica
    out._unit = crystal
    out._rep = (ra_n, rb_n, rc_n)
    return out


def strip_numbers_from_atom_labels(crystal: Crystal) -> None:
    """Truncate each species label at its first non-letter character
    (Xtals ``strip_numbers_from_atom_labels!``: "Zn1" -> "Zn")."""
    out = []
    for s in crystal.species:
        k = 0
        while k < len(s) and s[k].isalpha():
            k += 1
        out.append(s[:k])
    crystal.species = out
    crystal._device_cache.clear()


# ---------------------------------------------------------------- CIF ----
def _cif_number(tok: str) -> float:
    # strip standard uncertainties, e.g. 0.1234(5)
    return float(re.sub(r"\(.*\)$", "", tok))


def read_cif(path: str, species_col=("_atom_site_type_symbol", "_atom_site_label"),
             wrap_coords: bool = True) -> Crystal:
    """Read a P1 CIF into a Crystal (Xtals ``Crystal(filename)`` defaults)."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    cell = {}
    sym_ops = []
    loops = []
    i = 0
    n = len(lines)
    while i < n:
        ln = lines[i].strip()
        if not ln or ln.startswith("#"):
            i += 1
            continue
        if ln.startswith("loop_"):
            headers = []
            i += 1
            while i < n and lines[i].strip().startswith("_"):
                headers.append(lines[i].strip().split()[0])
                i += 1
            rows = []
            while i < n:
                s = lines[i].strip()
                if not s or s.startswith("_") or s.startswith("loop_") or s.startswith("data_") or s.startswith("#"):
                    break
                rows.append(shlex.split(s, posix=True))
                i += 1
            loops.append((headers, rows))
            continue
        if ln.startswith("_"):
            parts = shlex.split(ln, posix=True)
            key = parts[0]
            val = parts[1] if len(parts) > 1 else ""
            if key.startswith("_cell_"):
                cell[key] = _cif_number(val)
            if key in ("_symmetry_space_group_name_H-M", "_space_group_name_H-M_alt"):
                sym_ops.append(("hm", val.replace(" ", "")))
        i += 1
    a = cell["_cell_length_a"]
    b = cell["_cell_length_b"]
    c = cell["_cell_length_c"]
    deg = math.pi / 180.0
    box = Box(a, b, c, cell["_cell_angle_alpha"] * deg, cell["_cell_angle_beta"] * deg,
              cell["_cell_angle_gamma"] * deg)
    p1 = True
    for kind, val in sym_ops:
        if kind == "hm" and val.upper() not in ("P1", "P1(1)", "'P1'"):
            p1 = False
    atom_loop = None
    for headers, rows in loops:
        if "_atom_site_fract_x" in headers:
            atom_loop = (headers, rows)
        xyz_key = next((h for h in ("_symmetry_equiv_pos_as_xyz", "_space_group_symop_operation_xyz")
                        if h in headers), None)
        if xyz_key is not None:
            # the operation column only (a loop may also carry a site id column)
            k_op = headers.index(xyz_key)
            ops = [(" ".join(r[k_op:]) if len(r) != len(headers) else r[k_op]).replace(" ", "").lower()
                   for r in rows]
            if any(op not in ("x,y,z", "'x,y,z'") for op in ops):
                p1 = False
    if atom_loop is None:
        species, xf, q = [], np.zeros((3, 0)), np.zeros(0)
    else:
        headers, rows = atom_loop
        col = {h: k for k, h in enumerate(headers)}
        sp_key = next((k for k in species_col if k in col), None)
        if sp_key is None:
            raise ValueError(f"{path}: no species column among {species_col}")
        species = [r[col[sp_key]] for r in rows]
        xf = np.array([[_cif_number(r[col[k]]) for r in rows]
                       for k in ("_atom_site_fract_x", "_atom_site_fract_y", "_atom_site_fract_z")],
                      dtype=np.float64).reshape(3, -1)
        if "_atom_site_charge" in col:
            q = np.array([_cif_number(r[col["_atom_site_charge"]]) for r in rows], dtype=np.float64)
        else:
            q = np.zeros(len(rows))
    if wrap_coords:
        xf = np.mod(xf, 1.0)
    return Crystal(os.path.basename(path), box, species, xf, q, p1)
