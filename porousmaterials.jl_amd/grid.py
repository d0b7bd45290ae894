"""Energy grids with the reference's call surface, and the ``Grid`` type's
file format and index helpers.

``energy_grid(crystal, molecule, ljff; resolution, ...)`` follows
``src/grid.jl:291-362`` for the branch that runs in the reference, a
single-site probe: points at fractional (i, j, k)/(n - 1) of the unit cell
(0 and 1 included, ``:313``), the probe centred there in the replicated
crystal (``:326-338``), energies in K or kJ/mol (x 8.314/1000, ``:357-359``).
Data are a (n1, n2, n3) Fortran-ordered array, i fastest, like Julia's
``Array{Float64,3}``.

``write_cube`` / ``read_cube`` follow ``src/grid.jl:139-264`` (Gaussian-cube
text, no atoms, voxel vectors = lattice vectors / (n - 1), six values a line,
k fastest within (i, j)); ``xf_to_id`` / ``id_to_xf`` follow ``:34-61``.

A charged single-site probe (``ion``) gets the point-charge Ewald energy at
every point, the per-point ``electrostatic_potential_energy`` the reference's
tests use; the reference's own charged/multi-site branch does not run
(``src/grid.jl:342,347``) and multi-site probes raise here.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

from . import device
from .box import Box
from .crystal import assert_P1_symmetry, has_charges as xtal_has_charges, replicate
from .ewald import setup_Ewald_sum
from .forcefield import replication_factors
from .molecule import has_charges as mol_has_charges, needs_rotations


PATHS = {"grids": os.path.join(os.getcwd(), "data", "grids")}


def _isapprox(x, y, atol=0.0, rtol=None):
    """Julia ``isapprox`` of arrays: |x - y| <= max(atol, rtol max(|x|, |y|))
    in the 2-norm, rtol = sqrt(eps) unless atol > 0."""
    if rtol is None:
        rtol = 0.0 if atol > 0.0 else math.sqrt(np.finfo(np.float64).eps)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if x.shape != y.shape:
        return False
    d = float(np.linalg.norm((x - y).ravel()))
    return d <= max(atol, rtol * max(float(np.linalg.norm(x.ravel())), float(np.linalg.norm(y.ravel()))))


def box_isapprox(b1: Box, b2: Box, atol=0.0, rtol=None) -> bool:
    """Xtals ``isapprox(::Box, ::Box)``: the f_to_c matrices."""
    return _isapprox(b1.f_to_c, b2.f_to_c, atol, rtol)


@dataclass
class Grid:
    """``Grid{T}`` (src/grid.jl:12-18): data at the n_pts points of a regular
    fractional grid over ``box``, 0 and 1 included on every axis."""
    box: Box
    n_pts: tuple
    data: np.ndarray
    units: str
    origin: np.ndarray

    def isapprox(self, other, atol=0.0, rtol=None) -> bool:
        """``isapprox(g1, g2)`` (src/grid.jl:374-380)."""
        return (box_isapprox(self.box, other.box, atol, rtol) and tuple(self.n_pts) == tuple(other.n_pts)
                and _isapprox(self.data, other.data, atol, rtol) and self.units == other.units
                and _isapprox(self.origin, other.origin, atol, rtol))


def xf_to_id(n_pts, xf):
    """1-based voxel of fractional ``xf`` in an n_pts voxel partition of the
    unit cube, wrapped periodically (src/grid.jl:34-44)."""
    out = []
    for x, n in zip(xf, n_pts):
        v = math.floor(x * n) + 1
        if v <= 0:
            v += n
        elif v > n:
            v -= n
        out.append(v)
    return out


def id_to_xf(idx, n_pts):
    """Fractional coordinates of 1-based grid point ``idx`` (src/grid.jl:59-61)."""
    return [(i - 1.0) / (n - 1.0) for i, n in zip(idx, n_pts)]


def _fmt_e(v) -> str:
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Inf" if v > 0 else "-Inf"
    return "%e" % v


def write_cube(grid: Grid, filename: str, verbose=True, length_units="Angstrom") -> str:
    """Write ``grid`` to ``PATHS["grids"]/filename[.cube]``; returns the path."""
    if length_units not in ("Angstrom", "Bohr"):
        raise ValueError("length_units must be Angstrom or Bohr")
    os.makedirs(PATHS["grids"], exist_ok=True)
    if ".cube" not in filename:
        filename = filename + ".cube"
    path = os.path.join(PATHS["grids"], filename)
    n = tuple(int(x) for x in grid.n_pts)
    data = np.asarray(grid.data)
    lines = ["Units of data: %s\nLoop order: x, y, z\n" % grid.units,
             "%d %f %f %f\n" % (0, grid.origin[0], grid.origin[1], grid.origin[2])]
    for k in range(3):
        v = grid.box.f_to_c[:, k] / (n[k] - 1)
        if length_units == "Bohr":
            v = 1.88973 * v
        lines.append("%d %f %f %f\n" % (n[k], v[0], v[1], v[2]))
    for i in range(n[0]):
        for j in range(n[1]):
            row = []
            for k in range(n[2]):
                row.append(_fmt_e(data[i, j, k]) + " ")
                if (k + 1) % 6 == 0:
                    row.append("\n")
            row.append("\n")
            lines.append("".join(row))
    with open(path, "w") as f:
        f.write("".join(lines))
    if verbose:
        print("\tSee ", path)
    return path


def read_cube(filename: str, has_units=True) -> Grid:
    """Read a .cube file from ``PATHS["grids"]`` (atom lines skipped)."""
    if ".cube" not in filename:
        filename += ".cube"
    with open(os.path.join(PATHS["grids"], filename)) as f:
        text = f.read().split("\n")
    units = text[0].split()[3] if has_units else "missing"
    head = text[2].split()
    n_atoms = int(head[0])
    origin = np.array([float(x) for x in head[1:4]])
    n_pts, f_to_c = [], np.zeros((3, 3))
    for i in range(3):
        t = text[3 + i].split()
        n_pts.append(int(t[0]))
        f_to_c[:, i] = [float(x) for x in t[1:4]]
    for k in range(3):
        f_to_c[:, k] *= n_pts[k] - 1.0
    values = " ".join(text[6 + n_atoms:]).split()
    n = tuple(n_pts)
    total = n[0] * n[1] * n[2]
    if len(values) < total:
        raise ValueError(f"{filename}: {len(values)} values for a {n} grid")
    # file order: i outermost, k fastest
    data = np.array([float(v) for v in values[:total]]).reshape(n[0], n[1], n[2])
    return Grid(Box(f_to_c), n, np.asfortranarray(data), units, origin)


def required_n_pts(box: Box, dx: float):
    """``src/grid.jl:115-123``: ceil(|lattice vector| / dx) + 1 per axis."""
    norms = [float(np.linalg.norm(box.f_to_c[:, i])) for i in range(3)]
    return tuple(int(math.ceil(n / dx)) + 1 for n in norms)


def _grid_setup(crystal, molecule, ljff, resolution, units, center, verbose, ewald_precision,
                electrostatics=False):
    """Device handles of one grid, or ``None`` in place of them for a probe
    without LJ spheres when electrostatics are off (its grid is all zeros)."""
    assert_P1_symmetry(crystal)
    if units not in ("kJ_mol", "K"):
        raise ValueError("Pass kJ_mol or K for units of kJ/mol or K, respectively.")
    if molecule.frac:
        raise ValueError("energy_grid expects a Cartesian molecule")
    n_pts = tuple(int(x) for x in resolution) if isinstance(resolution, tuple) \
        else required_n_pts(crystal.box, float(resolution))
    if needs_rotations(molecule):
        raise NotImplementedError("energy grids of multi-site probes (Boltzmann average over "
                                  "rotations) are not supported")
    charged = electrostatics and xtal_has_charges(crystal) and mol_has_charges(molecule)
    grid_origin = crystal.box.f_to_c @ np.array([-0.5, -0.5, -0.5]) if center else np.zeros(3)
    if verbose:
        print(f"Computing energy grid of {molecule.species} in {crystal.name}")
        print(f"\tRegular grid (in fractional space) of {n_pts[0]} by {n_pts[1]} by {n_pts[2]} "
              f"points superimposed over the unit cell.")
    if molecule.n_atoms == 0 and not charged:
        # a single-site probe without an LJ sphere is a bare point charge: the
        # reference's grid is its vdw_energy, zero everywhere (src/grid.jl:337-338)
        return (None, None, None, None, n_pts), grid_origin
    repfactors = replication_factors(crystal.box, ljff)
    xtal = replicate(crystal, repfactors)
    fw = device.framework(xtal, ljff)
    ff = device.forcefield(ljff)
    ew = None
    if charged:
        eparams = setup_Ewald_sum(xtal.box, math.sqrt(ljff.r2_cutoff), eps=ewald_precision)
        ew = device.ewald(xtal, fw, eparams)
    pr = device.probe(molecule, ljff)
    return (fw, ff, ew, pr, n_pts), grid_origin


def _to_grid(crystal, n_pts, flat, units, origin):
    data = flat.reshape(n_pts, order="F")
    if units == "kJ_mol":
        data = data * 8.314 / 1000.0
    return Grid(crystal.box, n_pts, data, units, origin)


def energy_grid(crystal, molecule, ljff, resolution=1.0, n_rotations=1000, temperature=math.nan,
                units="kJ_mol", center=False, verbose=True, precision="fp64",
                ewald_precision=1e-6, electrostatics=False) -> Grid:
    """``src/grid.jl:291-362`` for single-site probes, on the device.

    As in the reference, the energy at a point is the probe's ``vdw_energy``
    alone (src/grid.jl:337-338), so a bare point charge (the only charged
    single-site probe) gives a grid of zeros.  ``electrostatics=True`` (no
    counterpart in the reference) adds the probe charge times the framework's
    Ewald potential, the point-charge energy of
    ``electrostatic_potential_energy(crystal, ion(q, xf), ...)``."""
    run, origin = _grid_setup(crystal, molecule, ljff, resolution, units, center, verbose,
                              ewald_precision, electrostatics)
    if run[3] is None:
        flat = np.zeros(int(np.prod(run[4])))
    else:
        flat = device.energy_grid(*run, precision)
    return _to_grid(crystal, run[4], flat, units, origin)


def energy_grids(crystals, molecule, ljff, resolution=1.0, units="kJ_mol", center=False,
                 verbose=False, precision="fp64", ewald_precision=1e-6, electrostatics=False):
    """``energy_grid`` of ``molecule`` in every crystal of ``crystals``, all
    grids in one device call (pm_energy_grid_batch).  Each grid equals what
    ``energy_grid`` returns for that crystal."""
    crystals = list(crystals)
    setups = [_grid_setup(x, molecule, ljff, resolution, units, center, verbose, ewald_precision, electrostatics)
              for x in crystals]
    live = [run for run, _ in setups if run[3] is not None]
    flats = iter(device.energy_grid_batch(live, precision) if live else [])
    return [_to_grid(x, run[4], next(flats) if run[3] is not None else np.zeros(int(np.prod(run[4]))),
                     units, origin)
            for x, (run, origin) in zip(crystals, setups)]
