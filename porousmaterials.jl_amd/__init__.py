"""MI355X-native guest-host energy engine with PorousMaterials.jl's call
surface for Widom insertions and energy grids.

The numerical work runs in libpmhip.so (hand-written gfx950 HIP kernels
behind the C ABI in include/pmhip.h); this package is the host-side mirror of
the reference's interface: ``henry_coefficient``, ``energy_grid``,
``vdw_energy``, ``electrostatic_potential_energy``,
``total_electrostatic_potential_energy`` and the data types they take.
"""
from .accessibility import accessible, compute_accessibility_grid
from .box import Box, replicate_box, unit_cube
from .crystal import (Crystal, assert_P1_symmetry, crystal_density, has_charges, net_charge,
                      read_cif, register_atomic_masses, replicate, strip_numbers_from_atom_labels)
from .energetics import (electrostatic_energies, electrostatic_potential_energy,
                         total_electrostatic_potential_energy, vdw_energies, vdw_energy)
from .ewald import (EPS0, FOUR_PI_EPS0, Eikr, EwaldParams, EwaldSum, determine_ewald_params,
                    make_eikr, precompute_kvec_wts, required_kreps, setup_Ewald_sum, total)
from .forcefield import LJForceField, forcefield_coverage, lennard_jones, replication_factors
from .energy_min import find_energy_minima, find_energy_minimum, find_energy_minimum_gridsearch
from .grid import (Grid, box_isapprox, energy_grid, energy_grids, id_to_xf, read_cube, required_n_pts,
                   write_cube, xf_to_id)
from .gcmc import (adsorption_isotherm, isotherm_sim_results_to_dataframe, muVT_sim, stepwise_adsorption_isotherm,
                   μVT_output_filename, μVT_sim)
from .henry import (K_TO_KJ_PER_MOL, N_BLOCKS, UNIV_GAS_CONST, block_statistics, henry_coefficient,
                    henry_coefficient_gradient, henry_coefficients,
                    henry_result_savename)
from .molecule import (Cart, Frac, Molecule, arvo_rotation_matrix, center_of_mass, ion, lj_sphere,
                       needs_rotations, point_charge, site_offsets, translate_to)
from .eos import PengRobinsonFluid, VdWFluid, calculate_properties
from . import device, eos, synthetic

__all__ = [n for n in dir() if not n.startswith("_")]
