"""Guest-host energies with the reference's call surface.

``vdw_energy(crystal, molecule, ljff)`` (``src/vdw_energetics.jl:62-64``),
``electrostatic_potential_energy(crystal, molecule, eparams, eikr)``
(``src/electrostatics.jl:331-342``) and
``total_electrostatic_potential_energy(crystal, molecules, eparams, eikr)``
(``src/electrostatics.jl:558-568``), evaluated by libpmhip.so.  As in the
reference the crystal must already be replicated enough for the nearest-image
convention (``replicate`` / ``replication_factors``), and fractional
coordinates refer to the crystal's box.
"""
from __future__ import annotations

import numpy as np

from . import device
from .ewald import EwaldSum
from .molecule import Frac

DEFAULT_PRECISION = "fp64"


def _frac(molecule, box):
    return molecule if molecule.frac else Frac(molecule, box)


def vdw_energy(crystal, molecule, ljff, precision=DEFAULT_PRECISION) -> float:
    """LJ energy (K) of one molecule pose; +Inf on overlap."""
    m = _frac(molecule, crystal.box)
    if m.n_atoms == 0:
        return 0.0
    fw = device.framework(crystal, ljff)
    ff = device.forcefield(ljff)
    sp = [ljff.index(s) for s in m.atom_species]
    return float(device.vdw_energies(fw, ff, sp, m.atoms_x.T[None, :, :], precision)[0])


def vdw_energies(crystal, site_species, site_xf, ljff, precision=DEFAULT_PRECISION) -> np.ndarray:
    """Batched form: LJ energy of each of ``site_xf.shape[0]`` poses of the same
    sites (``site_xf``: n_poses x n_sites x 3, fractional)."""
    fw = device.framework(crystal, ljff)
    ff = device.forcefield(ljff)
    sp = [ljff.index(s) for s in site_species]
    return device.vdw_energies(fw, ff, sp, site_xf, precision)


def _es_handles(crystal, eparams):
    fw = device.framework(crystal, None)
    return fw, device.ewald(crystal, fw, eparams)


def electrostatic_potential_energy(crystal, molecule, eparams, eikr=None,
                                   precision=DEFAULT_PRECISION) -> EwaldSum:
    """Ewald energy (K) of the molecule's point charges in the crystal's field."""
    m = _frac(molecule, crystal.box)
    if m.n_charges == 0 or crystal.charges_q.size == 0:
        return EwaldSum(0.0, 0.0)
    fw, ew = _es_handles(crystal, eparams)
    sr, lr = device.es_energies(fw, ew, m.charges_q, m.charges_x.T[None, :, :], precision)
    return EwaldSum(float(sr[0]), float(lr[0]))


def total_electrostatic_potential_energy(crystal, molecules, eparams, eikr=None,
                                         precision=DEFAULT_PRECISION) -> EwaldSum:
    """Sum over molecules; molecules with the same charges go in one batch."""
    out = EwaldSum(0.0, 0.0)
    if crystal.charges_q.size == 0:
        return out
    fw, ew = _es_handles(crystal, eparams)
    groups = {}
    for mol in molecules:
        m = _frac(mol, crystal.box)
        if m.n_charges == 0:
            continue
        groups.setdefault(tuple(m.charges_q.tolist()), []).append(m.charges_x.T)
    for q, xs in groups.items():
        sr, lr = device.es_energies(fw, ew, np.array(q), np.stack(xs), precision)
        out = out + EwaldSum(float(sr.sum()), float(lr.sum()))
    return out


def electrostatic_energies(crystal, charges, charge_xf, eparams, precision=DEFAULT_PRECISION):
    """Batched form: (sr, lr) arrays for n_poses x n_charges x 3 positions."""
    fw, ew = _es_handles(crystal, eparams)
    return device.es_energies(fw, ew, charges, charge_xf, precision)
