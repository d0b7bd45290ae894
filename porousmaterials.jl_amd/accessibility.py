"""Accessibility grids: which parts of a unit cell a probe can reach.

``compute_accessibility_grid(crystal, probe, forcefield; ...)`` follows
``src/grid.jl:689-774``: an energy grid of the probe (on the device), the
points below ``energy_tol`` and, with ``block_inaccessible_pockets``, only
those in segments through which the probe can leave the unit cell.  The
flood fill, the periodic segment graph and its simple cycles run on the host
inside libpmhip.so (pm_block_pockets), as the reference runs them on its host.

``accessible(grid, xf[, repfactors])`` follows ``src/grid.jl:778-836``: the
nearest grid point (``round``, ties to even) and its 26 neighbours, with the
reference's index wrap 0 -> n and n + 1 -> 1; a point is inaccessible only if
all 27 are.  ``henry_coefficient(...; accessibility_grid=grid)`` applies the
same test to every insertion on the device.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import device
from .crystal import assert_P1_symmetry
from .grid import Grid, energy_grid, write_cube


def _arg_nearest_neighbor(n_pts, xf):
    """1-based nearest grid point of fractional ``xf`` (src/grid.jl:778-780)."""
    return [1 + int(np.rint(float(x) * (int(n) - 1))) for x, n in zip(xf, n_pts)]


def _apply_pbc_to_index(idx, n_pts):
    """In place: index 0 -> n, n + 1 -> 1 (src/grid.jl:783-793)."""
    for k in range(3):
        if idx[k] == 0:
            idx[k] = n_pts[k]
        if idx[k] == n_pts[k] + 1:
            idx[k] = 1
    return idx


def accessible(accessibility_grid: Grid, xf, repfactors=None) -> bool:
    """Whether fractional ``xf`` of ``accessibility_grid.box`` (or of the box
    replicated by ``repfactors``, src/grid.jl:834-836) may be occupied."""
    if repfactors is not None:
        xf = [(float(x) * int(r)) % 1.0 for x, r in zip(xf, repfactors)]
    n_pts = tuple(int(n) for n in accessibility_grid.n_pts)
    data = np.asarray(accessibility_grid.data)
    nn = _arg_nearest_neighbor(n_pts, xf)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                idx = _apply_pbc_to_index([nn[0] + i, nn[1] + j, nn[2] + k], n_pts)
                if data[idx[0] - 1, idx[1] - 1, idx[2] - 1]:
                    return True
    return False


def _cube_name(probe, crystal, forcefield, stage):
    name = crystal.name.replace(".cif", "").replace(".cssr", "")
    return "%s_in_%s_%s_%s_pocket_blocking.cube" % (probe.species, name, forcefield.name.replace(".csv", ""),
                                                    stage)


def compute_accessibility_grid(crystal, probe, forcefield, resolution=1.0, energy_tol=10.0,
                               energy_units="kJ_mol", verbose=True, write_b4_after_grids=True,
                               block_inaccessible_pockets=True, precision="fp64", ewald_precision=1e-6):
    """(accessibility grid, number of segments blocked, porosity).

    Porosity is a float without pocket blocking and a dict
    {"b4_blocking", "after_blocking"} with it, as in the reference.  The
    energy grid is computed in ``precision`` (FP64 by default: the threshold
    then sees the energies the reference's FP64 code computes)."""
    assert_P1_symmetry(crystal)
    if verbose:
        print("Computing accessibility grid of %s using %f %s potential energy tol and %s probe..."
              % (crystal.name, energy_tol, energy_units, probe.species))
    grid = energy_grid(crystal, probe, forcefield, resolution=resolution, verbose=verbose, units=energy_units,
                       precision=precision, ewald_precision=ewald_precision)
    if not block_inaccessible_pockets:
        acc, _, _ = device.block_pockets(grid.data, energy_tol, block_pockets=False)
        agrid = Grid(grid.box, grid.n_pts, acc, "accessibility", grid.origin)
        return agrid, 0, float(acc.sum()) / acc.size

    acc, seg, (n_seg, n_blocked, n_cycles) = device.block_pockets(grid.data, energy_tol, block_pockets=True)
    if verbose:
        print("Found %d segments" % n_seg)
        print("\tFound %d simple cycles in segment connectivity graph." % n_cycles)
    if write_b4_after_grids:
        b4 = np.where(seg != -1, 1, -1).astype(np.int64)
        write_cube(Grid(grid.box, grid.n_pts, b4, "Segment_No", grid.origin),
                   _cube_name(probe, crystal, forcefield, "b4"), verbose=verbose)
        after = np.where(acc, 1, -1).astype(np.int64)
        write_cube(Grid(grid.box, grid.n_pts, after, "Segment_No", grid.origin),
                   _cube_name(probe, crystal, forcefield, "after"), verbose=verbose)
    if not acc.any():
        warnings.warn("%s cannot enter the pores of %s with %f K energy tolerance."
                      % (probe.species, crystal.name, energy_tol))
    porosity = {"b4_blocking": float(np.sum(grid.data < energy_tol)) / grid.data.size,
                "after_blocking": float(acc.sum()) / acc.size}
    if n_blocked != 0:  # printed whatever verbose says, as in the reference
        print("\t%d pockets in %s were found to be inaccessible to %s and blocked."
              % (n_blocked, crystal.name, probe.species))
        print("\tPorosity of %s b4 pocket blocking is %f" % (crystal.name, porosity["b4_blocking"]))
        print("\tPorosity of %s after pocket blocking is %f" % (crystal.name, porosity["after_blocking"]))
        print("\tpotential energy barrier used to determine accessibility: %f %s" % (energy_tol, energy_units))
    return Grid(grid.box, grid.n_pts, acc, "accessibility", grid.origin), n_blocked, porosity

