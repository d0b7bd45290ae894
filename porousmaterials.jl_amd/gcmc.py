"""Grand-canonical Monte Carlo with the reference's call surface.

``μVT_sim(xtal, molecule_templates, temperature, pressures, ljff; ...)``
follows src/gcmc.jl:175-817: cycle counts from the unit-cell volume
(``:200-205``), fugacities from an equation of state (``:228-243``),
replication for the cutoff (``:262-264``), the Markov chain (insertion,
deletion, translation with the adaptive step, rotation, reinsertion,
identity change; ``:404-627``), block statistics (``:629-655``), the energy
check (``:729-741``) and the results dict (``:748-807``).  The chain itself
runs on the device (libpmhip.so, ``pm_gcmc_*``): one workgroup per chain.

``adsorption_isotherm`` runs every pressure as its own chain in the same
device launches (the reference's ``pmap`` over pressures, src/gcmc.jl:996-
1027); ``stepwise_adsorption_isotherm`` runs them one after the other, each
starting from the previous end configuration (``:940-972``).

Differences from the reference, by design: the chain draws from a
counter-based Philox stream keyed by ``seed`` (the reference uses Julia's
global RNG), so a chain is reproducible and the oracle can replay it;
``n_burn_cycles`` / ``n_sample_cycles`` may be given directly (the older
keyword form the reference's long tests use); results are saved as JSON.
"""
from __future__ import annotations

import json
import math
import os
import time
import warnings

import numpy as np

from . import device
from .box import Box
from .crystal import ATOMIC_MASSES, assert_P1_symmetry, has_charges as xtal_has_charges, replicate
from .eos import PengRobinsonFluid, calculate_properties
from .forcefield import replication_factors
from .grid import Grid, required_n_pts
from .molecule import Molecule, has_charges as mol_has_charges

BOLTZMANN = 1.38064852e7  # Pa A^3 / K, src/PorousMaterials.jl:14
N_BLOCKS = 5
PROPOSAL_ENCODINGS = {1: "insertion", 2: "deletion", 3: "translation", 4: "rotation", 5: "reinsertion",
                      6: "identity change"}  # src/gcmc.jl:6-13
PATHS = {"simulations": os.path.join(os.getcwd(), "data", "simulations")}


def μVT_output_filename(xtal, molecule_templates, temperature, pressures, ljff, n_burn_cycles, n_sample_cycles,
                        comment="", extension=".json"):
    """src/gcmc.jl:830-839."""
    name = "muVT_xtal_%s_T_%.3fK" % (xtal.name, temperature)
    for m, p in zip(molecule_templates, pressures):
        name += "_%s_P_%.6fbar" % (m.species, p)
    name += "_%s_%dburn_%dsample" % (ljff.name, n_burn_cycles, n_sample_cycles)
    return name + comment + extension


def _fugacities(templates, temperature, pressures, eos):
    if eos == "ideal":
        return [p * 100000.0 for p in pressures]
    if eos == "PengRobinson":
        if len(templates) != 1:
            raise ValueError("Peng Robinson EOS not implemented for >1 species")
        props = calculate_properties(PengRobinsonFluid.from_table(templates[0].species), temperature, pressures[0],
                                     verbose=False)
        return [props["fugacity (bar)"] * 100000.0]
    raise ValueError("eos=:ideal and eos=:PengRobinson are the only valid options for an equation of state.")


def _cycles(xtal, n_cycles_per_volume, fraction_burn_cycles, n_burn_cycles, n_sample_cycles):
    if not (0.0 < fraction_burn_cycles < 1.0):
        raise ValueError("fraction_burn_cycles must lie in (0, 1)")
    nb = max(N_BLOCKS, math.ceil(n_cycles_per_volume * xtal.box.Omega))
    n_burn = math.ceil(nb * fraction_burn_cycles) if n_burn_cycles is None else int(n_burn_cycles)
    n_sample = math.ceil(nb * (1 - fraction_burn_cycles)) if n_sample_cycles is None else int(n_sample_cycles)
    if n_sample < N_BLOCKS:
        warnings.warn("# sample cycles set to minimum %d, which is number of blocks." % N_BLOCKS)
        n_sample = N_BLOCKS
    return n_burn, n_sample


def _check_inputs(xtal, templates, ljff):
    assert_P1_symmetry(xtal)
    for s in set(xtal.species):
        ljff.index(s)  # KeyError: not covered by the force field (src/gcmc.jl:364-366)
    for m in templates:
        if abs(float(np.sum(m.charges_q))) > 1e-6:
            raise ValueError("Molecule %s is not charge neutral!" % m.species)
        for s in m.atom_species:
            ljff.index(s)
        if m.n_atoms == 0:
            raise ValueError("molecule %s has no LJ sphere" % m.species)
    if xtal_has_charges(xtal) and any(mol_has_charges(m) for m in templates):
        raise ValueError("sorry, electrostatics not supported")  # src/gcmc.jl:378-382


def _molecular_weight(xtal):
    return float(sum(ATOMIC_MASSES[s] for s in xtal.species))


def _results(chain_state, xtal, xtal_rep, templates, temperature, pressures, fugacities, ljff, repfactors, n_burn,
             n_sample, elapsed, density_grid):
    """Statistics and the results dict of src/gcmc.jl:83-98,748-807."""
    st = chain_state
    S = len(templates)
    samples = st["block_samples"].astype(np.float64)
    n_sum = st["block_n"][:, :S].astype(np.float64)
    n2_sum = st["block_n2"][:, :S].astype(np.float64)
    tot_samples = samples.sum()
    avg_n = n_sum.sum(axis=0) / tot_samples
    avg_gh = st["block_gh"].sum() / tot_samples
    avg_gg = st["block_gg"].sum() / tot_samples
    with np.errstate(invalid="ignore", divide="ignore"):
        blk_n = n_sum / samples[:, None]
        blk_gh = st["block_gh"] / samples
        blk_gg = st["block_gg"] / samples
    err_n = 1.96 * np.std(blk_n, axis=0, ddof=1) / math.sqrt(N_BLOCKS)
    err_gh = 1.96 * float(np.std(blk_gh, ddof=1)) / math.sqrt(N_BLOCKS)
    err_gg = 1.96 * float(np.std(blk_gg, ddof=1)) / math.sqrt(N_BLOCKS)
    nr = int(np.prod(repfactors))
    r = {}
    r["xtal"] = xtal.name
    r["adsorbate"] = [m.species for m in templates]
    r["forcefield"] = ljff.name
    r["pressure (bar)"] = list(pressures)
    r["fugacity (bar)"] = [f / 100000.0 for f in fugacities]
    r["temperature (K)"] = temperature
    r["repfactors"] = tuple(int(x) for x in repfactors)
    r["# sample cycles"] = n_sample
    r["# burn cycles"] = n_burn
    r["# samples"] = int(tot_samples)
    r["elapsed time (min)"] = elapsed / 60.0
    r["⟨N⟩ (molecules)"] = avg_n
    r["⟨U_gh, vdw⟩ (K)"] = avg_gh
    r["⟨U_gh, electro⟩ (K)"] = 0.0
    r["⟨U_gg, vdw⟩ (K)"] = avg_gg
    r["⟨U_gg, electro⟩ (K)"] = 0.0
    r["⟨U⟩ (K)"] = avg_gg + 0.0 + avg_gh + 0.0
    r["var(N)"] = n2_sum.sum(axis=0) / tot_samples - avg_n ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        r["Q_st (K)"] = temperature - (st["block_un"].sum() / tot_samples - r["⟨U⟩ (K)"] * avg_n) / r["var(N)"]
    r["err ⟨N⟩ (molecules)"] = err_n
    r["err ⟨U_gh, vdw⟩ (K)"] = err_gh
    r["err ⟨U_gh, electro⟩ (K)"] = 0.0
    r["err ⟨U_gg, vdw⟩ (K)"] = err_gg
    r["err ⟨U_gg, electro⟩ (K)"] = 0.0
    r["err ⟨U⟩ (K)"] = err_gg + 0.0 + err_gh + 0.0
    r["⟨N⟩ (molecules/unit cell)"] = avg_n / nr
    r["err ⟨N⟩ (molecules/unit cell)"] = err_n / nr
    # an empty framework has no mass: Inf (NaN at zero loading), as Julia's float division gives
    with np.errstate(invalid="ignore", divide="ignore"):
        to_mmol_g = np.float64(1000.0) / (6.022140857e23 * _molecular_weight(xtal_rep) * 1.66054e-24) * nr
        r["⟨N⟩ (mmol/g)"] = r["⟨N⟩ (molecules/unit cell)"] * to_mmol_g
        r["err ⟨N⟩ (mmol/g)"] = r["err ⟨N⟩ (molecules/unit cell)"] * to_mmol_g
    for k, desc in PROPOSAL_ENCODINGS.items():
        prop, acc = int(st["n_proposed"][k - 1]), int(st["n_accepted"][k - 1])
        r["Total # %s proposals" % desc] = prop
        r["Fraction of %s proposals accepted" % desc] = acc / prop if prop else math.nan
    r["density grid"] = density_grid
    r["num snapshots"] = int(st["num_snapshots"])
    return r


def _jsonable(v):
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, Grid):
        return {"grid": True, "f_to_c": np.asarray(v.box.f_to_c).tolist(), "n_pts": list(v.n_pts), "units": v.units,
                "data": np.asarray(v.data).ravel(order="F").tolist(), "origin": np.asarray(v.origin).tolist()}
    if isinstance(v, (np.floating, np.integer)):
        return v.item()
    return v


def _from_json(key, v):
    """Inverse of ``_jsonable`` for a saved results dict."""
    if isinstance(v, dict) and v.get("grid"):
        return Grid(Box(np.array(v["f_to_c"])), tuple(v["n_pts"]),
                    np.array(v["data"], dtype=np.float64).reshape(tuple(v["n_pts"]), order="F"), v["units"],
                    np.array(v["origin"], dtype=np.float64))
    if key == "repfactors":
        return tuple(v)
    if isinstance(v, list) and all(isinstance(x, (int, float)) for x in v) and key not in ("pressure (bar)",
                                                                                           "fugacity (bar)"):
        return np.array(v)
    return v


def save_results(results, path):
    """Write a μVT_sim results dict as JSON (the reference writes JLD2)."""
    with open(path, "w") as f:
        json.dump({k: _jsonable(v) for k, v in results.items()}, f, indent=1, ensure_ascii=False)


def load_results(path):
    """Read a results dict written by ``save_results``."""
    with open(path) as f:
        return {k: _from_json(k, v) for k, v in json.load(f).items()}


def isotherm_sim_results_to_dataframe(desired_props, xtal, molecule_templates, temperature, pressures, ljff,
                                      n_burn_cycles, n_sample_cycles, comment="", where_are_jld_files=None):
    """``src/gcmc.jl:1081-1128``: a pandas DataFrame with one row per pressure
    set and the ``desired_props`` of each saved μVT_sim results dict (found by
    ``μVT_output_filename`` under ``PATHS["simulations"]`` or
    ``where_are_jld_files``)."""
    import pandas as pd

    if isinstance(molecule_templates, Molecule):
        molecule_templates = [molecule_templates]
        pressures = [[float(p)] for p in pressures]
    where = PATHS["simulations"] if where_are_jld_files is None else where_are_jld_files
    rows = []
    for p in pressures:
        name = μVT_output_filename(xtal, molecule_templates, temperature, list(p), ljff, n_burn_cycles,
                                   n_sample_cycles, comment=comment)
        res = load_results(os.path.join(where, name))
        rows.append({prop: res[prop] for prop in desired_props})
    return pd.DataFrame(rows, columns=list(desired_props))


def _run(xtal, templates, temperature, pressure_sets, ljff, molecule_sets, n_cycles_per_volume,
         fraction_burn_cycles, sample_frequency, verbose, eos, autosave, show_progress_bar, snapshot_frequency,
         calculate_density_grid, density_grid_dx, density_grid_molecular_species, density_grid_sim_box,
         results_filename_comment, n_burn_cycles, n_sample_cycles, seed, capacity, write_adsorbate_snapshots):
    """Chains for every pressure set in one pm_gcmc handle."""
    t0 = time.time()
    templates = list(templates)
    _check_inputs(xtal, templates, ljff)
    if write_adsorbate_snapshots:
        raise NotImplementedError("xyz snapshots of the adsorbates are not written by the device chains")
    for m in templates:
        if m.frac:
            raise ValueError("molecule templates must be Cartesian")
    S = len(templates)
    if any(len(p) != S for p in pressure_sets):
        raise ValueError("# molecules in simulation: %d; each pressure set needs as many partial pressures" % S)
    n_burn, n_sample = _cycles(xtal, n_cycles_per_volume, fraction_burn_cycles, n_burn_cycles, n_sample_cycles)
    fug_sets = [_fugacities(templates, temperature, p, eos) for p in pressure_sets]
    repfactors = replication_factors(xtal.box, ljff)
    xtal_rep = replicate(xtal, repfactors)
    box = xtal_rep.box
    # density grid (src/gcmc.jl:279-325)
    density = None
    species_names = [m.species for m in templates]
    if calculate_density_grid:
        if density_grid_molecular_species is None:
            if S != 1:
                raise ValueError("Pass `density_grid_molecular_species` to specify which adsorbate to make a grid for.")
            density_grid_molecular_species = species_names[0]
        sid = species_names.index(density_grid_molecular_species)
        dbox = box if density_grid_sim_box else xtal.box
        n_pts = required_n_pts(dbox, density_grid_dx)
        # update_density! counts the atoms whose species equals the molecule's species
        mask = sum(1 << k for k, a in enumerate(templates[sid].atom_species) if a == density_grid_molecular_species)
        density = {"species": sid, "site_mask": mask, "n_pts": n_pts, "sim_box": density_grid_sim_box}
    if capacity is None:
        # generous: the simulation box packed at 0.1 molecules per A^3
        capacity = int(min(1 << 18, max(256, math.ceil(0.1 * box.Omega))))
    fw = device.framework(xtal_rep, ljff)
    chains = device.GcmcChains(fw, ljff, templates, [temperature] * len(pressure_sets), fug_sets, n_burn, n_sample,
                               sample_frequency=sample_frequency, snapshot_frequency=snapshot_frequency,
                               capacity=capacity, density=density, seed=seed)
    for c, mols in enumerate(molecule_sets):
        if mols is None:
            continue
        for s, mol_list in enumerate(mols):
            if not mol_list:
                continue
            sites, coms = [], []
            for mol in mol_list:
                if mol.species != species_names[s]:
                    raise ValueError("initializing with wrong molecule species")
                mf = mol if mol.frac else _to_frac(mol, box)
                sites.append(mf.atoms_x.T)
                coms.append(mf.com)
            chains.set_molecules(c, s, np.array(sites), np.array(coms))
    if verbose:
        print("μVT simulation of %s in %s at %f K (%d chains), %d burn + %d sample cycles"
              % (", ".join(species_names), xtal.name, temperature, len(pressure_sets), n_burn, n_sample))
    total = n_burn + n_sample
    done, step = 0, 1
    last_print = time.time()
    while done < total:
        k = min(step, total - done)
        ts = time.time()
        chains.run(k)
        dt = time.time() - ts
        done += k
        # grow the launch towards ~1 s of device time
        step = max(1, min(total, int(k * max(0.5, min(4.0, 1.0 / max(dt, 1e-3))))))
        if (verbose or show_progress_bar) and time.time() - last_print > 10.0:
            print("\tcycle %d / %d" % (done, total), flush=True)
            last_print = time.time()
    elapsed = time.time() - t0
    out = []
    for c, (pressures, fugs) in enumerate(zip(pressure_sets, fug_sets)):
        st = chains.state(c)
        if st["error"]:
            raise RuntimeError("chain %d reached the molecule capacity %d; pass a larger capacity" % (c, capacity))
        gh, gg = chains.total_energy(c)
        if not (abs(gh - st["energy_gh"]) <= 0.01 and abs(gg - st["energy_gg"]) <= 0.01):
            raise RuntimeError("energy incremented improperly during simulation...")
        dgrid = Grid(box if density_grid_sim_box else xtal.box, (0, 0, 0), np.zeros((0, 0, 0)), "inverse_A3",
                     np.zeros(3))
        if density is not None:
            counts = chains.density(c)
            dgrid = Grid(box if density_grid_sim_box else xtal.box, tuple(density["n_pts"]),
                         counts / max(st["num_snapshots"], 1), "inverse_A3", np.zeros(3))
        res = _results(st, xtal, xtal_rep, templates, temperature, pressures, fugs, ljff, repfactors, n_burn,
                       n_sample, elapsed, dgrid)
        mols = []
        for s in range(S):
            sites, coms = chains.molecules(c, s)
            mols.append([_device_molecule(templates[s], sites[i], coms[i], box) for i in range(sites.shape[0])])
        if verbose:
            print_results(res)
        if autosave:
            os.makedirs(PATHS["simulations"], exist_ok=True)
            name = os.path.join(PATHS["simulations"],
                                μVT_output_filename(xtal, templates, temperature, pressures, ljff, n_burn, n_sample,
                                                    comment=results_filename_comment))
            save_results(res, name)
            if verbose:
                print("\tresults dictionary saved in ", name)
        out.append((res, mols))
    return out


def _device_molecule(tmpl, sites_xf, com_xf, box):
    """A device molecule as the reference returns it (Cartesian in the
    simulation box); the charges follow the rigid motion that takes the
    template's LJ spheres onto the device's (Kabsch)."""
    x = box.f_to_c @ np.asarray(sites_xf).T
    c = box.f_to_c @ np.asarray(com_xf)
    P = tmpl.atoms_x - tmpl.com[:, None]
    U, _, Vt = np.linalg.svd((x - c[:, None]) @ P.T)
    R = U @ np.diag([1.0, 1.0, np.sign(np.linalg.det(U @ Vt))]) @ Vt
    qx = c[:, None] + R @ (tmpl.charges_x - tmpl.com[:, None])
    return Molecule(tmpl.species, list(tmpl.atom_species), x, np.array(tmpl.charges_q, dtype=np.float64), qx, c)


def _to_frac(mol, box):
    out = mol.copy()
    out.atoms_x = box.c_to_f @ mol.atoms_x
    out.charges_x = box.c_to_f @ mol.charges_x
    out.com = box.c_to_f @ mol.com
    out.frac = True
    return out


def print_results(results):
    """src/gcmc.jl:842-892 (abridged)."""
    print("Unit cell replication factors: %d %d %d" % tuple(results["repfactors"]))
    for key in ["# sample cycles", "# burn cycles", "# samples"]:
        print(key + ": ", results[key])
    for desc in PROPOSAL_ENCODINGS.values():
        total = results["Total # %s proposals" % desc]
        if total > 0:
            print("%s\t%d total proposals.\n\t%f %% proposals accepted." %
                  (desc, total, 100.0 * results["Fraction of %s proposals accepted" % desc]))
    for s, name in enumerate(results["adsorbate"]):
        print(name)
        for key in ["⟨N⟩ (molecules)", "⟨N⟩ (molecules/unit cell)", "⟨N⟩ (mmol/g)"]:
            print("%s: %f +/- %f" % (key, results[key][s], results["err " + key][s]))
    for key in ["⟨U_gg, vdw⟩ (K)", "⟨U_gh, vdw⟩ (K)", "⟨U_gg, electro⟩ (K)", "⟨U_gh, electro⟩ (K)", "⟨U⟩ (K)"]:
        print("%s: %f +/- %f" % (key, results[key], results["err " + key]))


_DEFAULTS = dict(n_cycles_per_volume=200, fraction_burn_cycles=0.5, sample_frequency=1, verbose=True,
                 ewald_precision=1e-6, eos="ideal", autosave=True, show_progress_bar=False,
                 write_adsorbate_snapshots=False, snapshot_frequency=1, calculate_density_grid=False,
                 density_grid_dx=1.0, density_grid_molecular_species=None, density_grid_sim_box=True,
                 results_filename_comment="", n_burn_cycles=None, n_sample_cycles=None, seed=0, capacity=None)


def _kw(kwargs):
    unknown = set(kwargs) - set(_DEFAULTS)
    if unknown:
        raise TypeError("unexpected keyword arguments %s" % sorted(unknown))
    out = dict(_DEFAULTS)
    out.update(kwargs)
    out.pop("ewald_precision")  # the reference's GCMC has no electrostatics
    return out


def μVT_sim(xtal, molecule_templates, temperature, pressures, ljff, molecules=None, **kwargs):
    """(results, molecules) of one grand-canonical simulation
    (src/gcmc.jl:175-817).  ``molecule_templates`` and ``pressures`` may be a
    single Molecule and a float (src/gcmc.jl:820)."""
    if isinstance(molecule_templates, Molecule):
        molecule_templates, pressures = [molecule_templates], [float(pressures)]
    kw = _kw(kwargs)
    return _run(xtal, molecule_templates, temperature, [list(pressures)], ljff, [molecules], **kw)[0]


muVT_sim = μVT_sim


def adsorption_isotherm(xtal, molecule_templates, temperature, pressures, ljff, **kwargs):
    """Results dicts of μVT_sim at each pressure (src/gcmc.jl:996-1027), every
    pressure a separate chain of one device handle."""
    if isinstance(molecule_templates, Molecule):
        molecule_templates = [molecule_templates]
        pressures = [[float(p)] for p in pressures]
    kw = _kw(kwargs)
    out = _run(xtal, molecule_templates, temperature, [list(p) for p in pressures], ljff,
               [None] * len(pressures), **kw)
    return [r for r, _ in out]


def stepwise_adsorption_isotherm(xtal, molecule_templates, temperature, pressures, ljff, **kwargs):
    """Results dicts of μVT_sim at each pressure in turn, each simulation
    starting from the previous one's molecules (src/gcmc.jl:940-972)."""
    if isinstance(molecule_templates, Molecule):
        molecule_templates = [molecule_templates]
        pressures = [[float(p)] for p in pressures]
    if len(molecule_templates) != 1:
        raise ValueError("We only calculate pure gas isotherms")
    kw = _kw(kwargs)
    seed = kw.pop("seed")
    results, molecules = [], None
    for i, p in enumerate(pressures):
        # each step its own stream
        r, molecules = _run(xtal, molecule_templates, temperature, [list(p)], ljff, [molecules], seed=seed + i,
                            **kw)[0]
        results.append(r)
    return results
