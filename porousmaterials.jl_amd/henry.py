"""Widom-insertion Henry coefficient with the reference's call surface.

``henry_coefficient(crystal, molecule, temperature, ljff; ...)`` follows
``src/henry.jl:34-181``: insertion count from the unit-cell volume (``:67``),
``N_BLOCKS = 5`` equal blocks (``:77``, ``src/mc_helpers.jl:11``),
replication for the cutoff (``:80-86``), per-block Boltzmann sums computed by
libpmhip.so (``:185-265``), per-block K_H and <U> and their 2 sigma/sqrt(5)
errors (``:112-157``) and the same 16-key results dict.

Differences from the reference, by design:

* insertions come from a counter-based Philox4x32-10 stream keyed by
  ``seed`` (the reference draws from the unseeded global RNG, ``:214``), so a
  run is reproducible and the oracle can replay every insertion;
* each insertion rotates the template by a fresh Arvo rotation (the reference
  composes rotations cumulatively, ``:218-220``; both are uniform on SO(3)),
  so there is no bond-length drift to check (``:247-261``);
* all blocks run in one device launch instead of a ``pmap`` over workers.

With ``accessibility_grid`` (a ``Grid`` of booleans over the crystal's unit
cell, from ``compute_accessibility_grid``) an insertion whose centre of mass
is not ``accessible(grid, xf, repfactors)`` counts with E = +Inf and is not
computed (``:225-232``); as in the reference, the grid's box is checked
against the crystal's only when ``verbose`` (``:57-63``).
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import device
from .crystal import assert_P1_symmetry, crystal_density, has_charges as xtal_has_charges, replicate
from .ewald import setup_Ewald_sum
from .forcefield import replication_factors
from .grid import box_isapprox
from .molecule import has_charges as mol_has_charges

UNIV_GAS_CONST = 8.3144598e-5    # m^3 bar / (K mol), src/PorousMaterials.jl:12
K_TO_KJ_PER_MOL = 8.3144598e-3   # kJ / (mol K), src/PorousMaterials.jl:13
N_BLOCKS = 5                     # src/mc_helpers.jl:11

PATHS = {"simulations": os.path.join(os.getcwd(), "data", "simulations")}


def henry_result_savename(crystal, molecule, temperature, ljff, insertions_per_volume, comment=""):
    if comment != "" and comment[0] != "_":
        comment = "_" + comment
    return "henry_sim_%s_in_%s_%fK_%s_ff_%d_insertions_per_volume%s.json" % (
        molecule.species, crystal.name, temperature, ljff.name, int(insertions_per_volume), comment)


def block_statistics(block_sums, nb_per_block, temperature, rho):
    """Per-block K_H and <U> and the results dict of src/henry.jl:112-159.
    block_sums: (N_BLOCKS, 3) of (sum w, sum w E_vdw, sum w E_es)."""
    s = block_sums[:, 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        u_vdw = block_sums[:, 1] / s
        u_es = block_sums[:, 2] / s
    kh = s / (UNIV_GAS_CONST * temperature * nb_per_block)  # mol/(m^3 bar)
    nb = len(s)
    err_kh = 2.0 * np.std(kh, ddof=1) / math.sqrt(nb)
    err_vdw = 2.0 * np.std(u_vdw, ddof=1) / math.sqrt(nb)
    err_es = 2.0 * np.std(u_es, ddof=1) / math.sqrt(nb)
    r = {}
    r["henry coefficient [mol/(m³-bar)]"] = float(np.mean(kh))
    r["henry coefficient [mmol/(g-bar)]"] = r["henry coefficient [mol/(m³-bar)]"] / rho
    r["err henry coefficient [mmol/(g-bar)]"] = float(err_kh / rho)
    r["henry coefficient [mol/(kg-Pa)]"] = r["henry coefficient [mmol/(g-bar)]"] / 100000.0
    r["⟨U, vdw⟩ (K)"] = float(np.mean(u_vdw))
    r["⟨U, es⟩ (K)"] = float(np.mean(u_es))
    r["⟨U⟩ (K)"] = r["⟨U, vdw⟩ (K)"] + r["⟨U, es⟩ (K)"]
    r["⟨U⟩ (kJ/mol)"] = r["⟨U⟩ (K)"] * K_TO_KJ_PER_MOL
    r["⟨U, vdw⟩ (kJ/mol)"] = r["⟨U, vdw⟩ (K)"] * K_TO_KJ_PER_MOL
    r["err ⟨U, vdw⟩ (kJ/mol)"] = float(err_vdw * K_TO_KJ_PER_MOL)
    r["⟨U, es⟩ (kJ/mol)"] = r["⟨U, es⟩ (K)"] * K_TO_KJ_PER_MOL
    r["err ⟨U, es⟩ (kJ/mol)"] = float(err_es * K_TO_KJ_PER_MOL)
    r["Qst (kJ/mol)"] = -r["⟨U⟩ (kJ/mol)"] + temperature * K_TO_KJ_PER_MOL
    r["err Qst (kJ/mol)"] = float((err_vdw + err_es) * K_TO_KJ_PER_MOL)
    return r


def _access_handle(crystal, accessibility_grid, verbose):
    if accessibility_grid is None:
        return None
    if verbose:
        print("Using provided accessibility grid to block inaccessible pockets")
        if not box_isapprox(accessibility_grid.box, crystal.box):
            raise ValueError("accessibility grid box does not match box of %s.\n" % crystal.name)
    return device.access(accessibility_grid)


def _widom_setup(crystal, molecule, temperature, ljff, insertions_per_volume, ewald_precision, verbose):
    """Insertion count, replication and device handles of one Henry run
    (src/henry.jl:41-109)."""
    assert_P1_symmetry(crystal)
    if molecule.frac:
        raise ValueError("henry_coefficient expects a Cartesian molecule template")
    nb_insertions = max(N_BLOCKS, math.ceil(insertions_per_volume * crystal.box.Omega))
    nb_per_block = math.ceil(nb_insertions / N_BLOCKS)
    repfactors = replication_factors(crystal.box, ljff)
    xtal = replicate(crystal, repfactors)
    charged = xtal_has_charges(crystal) and mol_has_charges(molecule)
    rho = crystal_density(xtal)
    if verbose:
        print(f"Simulating Henry coefficient of {molecule.species} in {crystal.name} at "
              f"{temperature} K, using {ljff.name} force field with {insertions_per_volume} "
              f"insertions per A^3.")
        print(f"\t{nb_insertions} total # particle insertions")
        print(f"\tReplicating crystal {repfactors[0]} by {repfactors[1]} by {repfactors[2]} "
              f"for short-range cutoff {math.sqrt(ljff.r2_cutoff):.2f}")
    fw = device.framework(xtal, ljff)
    ff = device.forcefield(ljff)
    ew = None
    if charged:
        eparams = setup_Ewald_sum(xtal.box, math.sqrt(ljff.r2_cutoff), eps=ewald_precision,
                                  verbose=verbose)
        ew = device.ewald(xtal, fw, eparams)
    pr = device.probe(molecule, ljff)
    return fw, ff, ew, pr, nb_per_block, rho


def _finish(results, crystal, molecule, temperature, ljff, insertions_per_volume, sums, nb_per_block,
            rho, t0, verbose, autosave, filename_comment):
    results["xtal"] = crystal.name
    results["elapsed time (min)"] = (time.time() - t0) / 60.0
    if verbose:
        for b in range(N_BLOCKS):
            print(f"\tBlock  {b + 1}/{N_BLOCKS} statistics:")
            print(f"\tHenry coeff. [mmol/(g-bar)]: "
                  f"{sums[b, 0] / (UNIV_GAS_CONST * temperature * nb_per_block) / rho}")
        print("\t----- final results ----")
        for key in ["henry coefficient [mmol/(g-bar)]", "⟨U, vdw⟩ (kJ/mol)", "⟨U, es⟩ (kJ/mol)",
                    "Qst (kJ/mol)"]:
            print(f"\t{key} = {results[key]:f} +/- {results['err ' + key]:f}")
    if autosave:
        os.makedirs(PATHS["simulations"], exist_ok=True)
        name = os.path.join(PATHS["simulations"],
                            henry_result_savename(crystal, molecule, temperature, ljff,
                                                  insertions_per_volume, filename_comment))
        with open(name, "w") as f:
            json.dump(results, f, indent=1, ensure_ascii=False)
        if verbose:
            print("\tResults saved in: ", name)
    return results


def henry_coefficient(crystal, molecule, temperature, ljff, insertions_per_volume=200,
                      verbose=True, ewald_precision=1e-6, autosave=True, filename_comment="",
                      accessibility_grid=None, seed=0, precision="fp64", return_block_sums=False):
    """K_H and <U> of ``molecule`` (Cartesian template) in ``crystal``."""
    t0 = time.time()
    acc = _access_handle(crystal, accessibility_grid, verbose)
    fw, ff, ew, pr, nb_per_block, rho = _widom_setup(crystal, molecule, temperature, ljff,
                                                     insertions_per_volume, ewald_precision, verbose)
    sums = device.widom(fw, ff, ew, pr, temperature, seed, N_BLOCKS, nb_per_block, precision, acc=acc)
    results = block_statistics(sums, nb_per_block, temperature, rho)
    _finish(results, crystal, molecule, temperature, ljff, insertions_per_volume, sums, nb_per_block,
            rho, t0, verbose, autosave, filename_comment)
    if return_block_sums:
        return results, sums
    return results


def henry_coefficients(pairs, temperature, ljff, insertions_per_volume=200, verbose=False,
                       ewald_precision=1e-6, autosave=False, filename_comment="", seeds=None,
                       precision="fp64", return_block_sums=False, accessibility_grids=None):
    """Screening: ``henry_coefficient`` for every (crystal, molecule) in
    ``pairs``, all runs in one device call (pm_widom_batch).  Each result is
    what ``henry_coefficient`` returns for that pair with the same seed
    (``seeds[i]``, default 0) and accessibility grid
    (``accessibility_grids[i]``, default None).  Returns the list of results
    dicts."""
    t0 = time.time()
    pairs = list(pairs)
    seeds = [0] * len(pairs) if seeds is None else list(seeds)
    grids = [None] * len(pairs) if accessibility_grids is None else list(accessibility_grids)
    if len(seeds) != len(pairs) or len(grids) != len(pairs):
        raise ValueError("one seed and one accessibility grid (or None) per (crystal, molecule) pair")
    setups, runs = [], []
    for (crystal, molecule), seed, agrid in zip(pairs, seeds, grids):
        acc = _access_handle(crystal, agrid, verbose)
        fw, ff, ew, pr, nb_per_block, rho = _widom_setup(crystal, molecule, temperature, ljff,
                                                         insertions_per_volume, ewald_precision,
                                                         verbose)
        setups.append((nb_per_block, rho))
        runs.append((fw, ff, ew, pr, temperature, seed, N_BLOCKS, nb_per_block, acc))
    all_sums = device.widom_batch(runs, precision)
    out = []
    for (crystal, molecule), (nb_per_block, rho), sums in zip(pairs, setups, all_sums):
        results = block_statistics(sums, nb_per_block, temperature, rho)
        _finish(results, crystal, molecule, temperature, ljff, insertions_per_volume, sums,
                nb_per_block, rho, t0, verbose, autosave, filename_comment)
        out.append((results, sums) if return_block_sums else results)
    return out


def henry_coefficient_gradient(crystal, molecule, temperature, ljff, insertions_per_volume=200,
                               ewald_precision=1e-6, seed=0, precision="fp64", verbose=False,
                               accessibility_grid=None):
    """K_H (the ``henry_coefficient`` results dict) and its logarithmic
    derivatives with respect to the framework species' pure LJ parameters and
    a uniform shift of their charges, from one adjoint device pass
    (pm_widom_gradient):

        d ln K_H / d theta = -(1/T) sum_b sum_{i in b} w_i dE_i/d theta / sum_b S_b

    (K_H = mean_b S_b / (R T n_b), ``src/henry.jl:112-135``, differentiated).
    Returns (results, grad) with grad[species] = {"epsilon": ..., "sigma": ...,
    "charge": ...} for every species in the framework; epsilon and sigma are
    the framework side's parameters, the probe's held fixed."""
    t0 = time.time()
    acc = _access_handle(crystal, accessibility_grid, verbose)
    fw, ff, ew, pr, nb_per_block, rho = _widom_setup(crystal, molecule, temperature, ljff,
                                                     insertions_per_volume, ewald_precision, verbose)
    sums, grads = device.widom_gradient(fw, ff, ew, pr, temperature, seed, N_BLOCKS, nb_per_block,
                                        precision, ljff, acc=acc)
    results = block_statistics(sums, nb_per_block, temperature, rho)
    results["xtal"] = crystal.name
    results["elapsed time (min)"] = (time.time() - t0) / 60.0
    s_tot = float(np.sum(sums[:, 0]))
    g_tot = grads.sum(axis=0)
    grad = {}
    for sp in sorted(set(crystal.species), key=ljff.index):
        k = ljff.index(sp)
        d = -g_tot[k] / (temperature * s_tot)
        grad[sp] = {"epsilon": float(d[0]), "sigma": float(d[1]), "charge": float(d[2])}
    return results, grad
