"""Energy grids with the reference's call surface.

``energy_grid(crystal, molecule, ljff; resolution, ...)`` follows
``src/grid.jl:291-362`` for the branch that runs in the reference, a
single-site probe: points at fractional (i, j, k)/(n - 1) of the unit cell
(0 and 1 included, ``:313``), the probe centred there in the replicated
crystal (``:326-338``), energies in K or kJ/mol (x 8.314/1000, ``:357-359``).
Data are a (n1, n2, n3) Fortran-ordered array, i fastest, like Julia's
``Array{Float64,3}``.

A charged single-site probe (``ion``) gets the point-charge Ewald energy at
every point, the per-point ``electrostatic_potential_energy`` the reference's
tests use; the reference's own charged/multi-site branch does not run
(``src/grid.jl:342,347``) and multi-site probes raise here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import device
from .box import Box
from .crystal import assert_P1_symmetry, has_charges as xtal_has_charges, replicate
from .ewald import setup_Ewald_sum
from .forcefield import replication_factors
from .molecule import has_charges as mol_has_charges, needs_rotations


@dataclass
class Grid:
    box: Box
    n_pts: tuple
    data: np.ndarray
    units: str
    origin: np.ndarray


def required_n_pts(box: Box, dx: float):
    """``src/grid.jl:115-123``: ceil(|lattice vector| / dx) + 1 per axis."""
    norms = [float(np.linalg.norm(box.f_to_c[:, i])) for i in range(3)]
    return tuple(int(math.ceil(n / dx)) + 1 for n in norms)


def _grid_setup(crystal, molecule, ljff, resolution, units, center, verbose, ewald_precision):
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
    charged = xtal_has_charges(crystal) and mol_has_charges(molecule)
    grid_origin = crystal.box.f_to_c @ np.array([-0.5, -0.5, -0.5]) if center else np.zeros(3)
    if verbose:
        print(f"Computing energy grid of {molecule.species} in {crystal.name}")
        print(f"\tRegular grid (in fractional space) of {n_pts[0]} by {n_pts[1]} by {n_pts[2]} "
              f"points superimposed over the unit cell.")
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
                units="kJ_mol", center=False, verbose=True, precision="mixed",
                ewald_precision=1e-6) -> Grid:
    run, origin = _grid_setup(crystal, molecule, ljff, resolution, units, center, verbose,
                              ewald_precision)
    flat = device.energy_grid(*run, precision)
    return _to_grid(crystal, run[4], flat, units, origin)


def energy_grids(crystals, molecule, ljff, resolution=1.0, units="kJ_mol", center=False,
                 verbose=False, precision="mixed", ewald_precision=1e-6):
    """``energy_grid`` of ``molecule`` in every crystal of ``crystals``, all
    grids in one device call (pm_energy_grid_batch).  Each grid equals what
    ``energy_grid`` returns for that crystal."""
    crystals = list(crystals)
    setups = [_grid_setup(x, molecule, ljff, resolution, units, center, verbose, ewald_precision)
              for x in crystals]
    flats = device.energy_grid_batch([run for run, _ in setups], precision)
    return [_to_grid(x, run[4], flat, units, origin)
            for x, (run, origin), flat in zip(crystals, setups, flats)]
