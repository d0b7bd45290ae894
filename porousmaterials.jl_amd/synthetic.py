"""Seeded synthetic frameworks and probes for the benchmark workloads.

Every case is drawn from ``numpy.random.Generator(PCG64(seed))`` so the same
seed gives the same framework on any machine.  Framework atoms are placed
uniformly in fractional space by sequential rejection: a candidate is kept if
its nearest-image Cartesian distance to every atom already placed is at least
``min_sep`` (1.0 A).  Framework LJ species are ``T1..Tn`` with epsilon ~
U[20, 80] K and sigma ~ U[2.5, 3.8] A drawn per species; charges ~
U[-1.5, 1.5] e shifted to zero net charge.

Workloads (sizes are the defaults of each constructor):

* ``lj_grid_case``       cubic a = 25.8 A, 424 atoms of 4 species, no
  charges, methane probe (148 K, 3.73 A), r_c 12.8 A, 50^3 grid.
* ``co2_widom_case``     triclinic a, b, c ~ U[18, 22] A, angles ~ U[80, 100]
  deg, 600 atoms of 5 species with charges, rigid 3-site CO2 (C-O 1.16 A,
  C +0.70 e 27 K 2.80 A, O -0.35 e 79 K 3.05 A), 298 K, Ewald 1e-6, r_c 12.8 A,
  10^6 insertions = 5 blocks x 200,000.
* ``screening_cases``    256 triclinic frameworks, N ~ logU[50, 3000],
  lengths ~ U[8, 40] A, angles ~ U[70, 110] deg.  Density rule: N is capped
  at floor(0.05 atoms/A^3 x V) (at least 1), otherwise random packing of up to
  3000 atoms into an 8 A cell would overlap every insertion.
* ``large_grid_case``    N = 3000, lengths ~ U[30, 40] A, angles ~ U[75, 105]
  deg, charges as above, 200^3 grid, unit point-charge and methane probes.
* ``small_grid_cases``   32 frameworks, N ~ U[200, 1500], lengths
  (N / 0.05)^(1/3) x U[0.9, 1.1] A, angles ~ U[75, 105] deg, 80^3 grids.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .box import Box
from .crystal import Crystal, register_atomic_masses
from .forcefield import LJForceField
from .molecule import Molecule, lj_sphere, point_charge

DEG = math.pi / 180.0
SPECIES_MASSES = [12.0107, 14.0067, 15.9994, 30.973762, 65.38, 26.9815386, 28.0855, 32.065]
PROBE_PARAMS = {"CH4": (3.73, 148.0), "C_CO2": (2.80, 27.0), "O_CO2": (3.05, 79.0)}


@dataclass
class Case:
    name: str
    crystal: Crystal
    ljff: LJForceField
    probes: dict
    temperature: float = 298.0
    ewald_precision: float = 1e-6
    meta: dict = field(default_factory=dict)


def place_atoms(rng, box: Box, n: int, min_sep: float = 1.0, max_tries: int = 2000000) -> np.ndarray:
    """(3, n) fractional coordinates, nearest-image separation >= min_sep."""
    pts = np.empty((n, 3))
    f = box.f_to_c
    k = 0
    tries = 0
    m2 = min_sep * min_sep
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n} atoms with separation {min_sep} A")
        c = rng.random(3)
        if k:
            d = pts[:k] - c
            d -= np.round(d)
            x = d @ f.T
            if np.min(np.einsum("ij,ij->i", x, x)) < m2:
                continue
        pts[k] = c
        k += 1
    return pts.T.copy()


def _species_names(n):
    return [f"T{i + 1}" for i in range(n)]


def _register(n):
    register_atomic_masses({f"T{i + 1}": SPECIES_MASSES[i % len(SPECIES_MASSES)] for i in range(n)})


def _forcefield(rng, n_species, r_cutoff, name):
    names = _species_names(n_species)
    eps = rng.uniform(20.0, 80.0, n_species)
    sig = rng.uniform(2.5, 3.8, n_species)
    pure_s = {s: float(v) for s, v in zip(names, sig)}
    pure_e = {s: float(v) for s, v in zip(names, eps)}
    for s, (sg, ep) in PROBE_PARAMS.items():
        pure_s[s] = sg
        pure_e[s] = ep
    return LJForceField(name, pure_s, pure_e, r_cutoff, "Lorentz-Berthelot")


def _framework(rng, name, box, n_atoms, n_species, charged):
    _register(n_species)
    xf = place_atoms(rng, box, n_atoms)
    types = rng.integers(0, n_species, n_atoms)
    species = [f"T{t + 1}" for t in types]
    q = None
    if charged:
        q = rng.uniform(-1.5, 1.5, n_atoms)
        q = q - q.mean()
    return Crystal(name, box, species, xf, q)


def co2_molecule(bond: float = 1.16) -> Molecule:
    ax = np.array([[0.0, -bond, bond], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    return Molecule("CO2", ["C_CO2", "O_CO2", "O_CO2"], ax, [0.70, -0.35, -0.35], ax.copy(),
                    np.zeros(3))


def lj_grid_case(seed: int = 11, n_atoms: int = 424, a: float = 25.8, n_species: int = 4,
                 n_pts: int = 50) -> Case:
    rng = np.random.Generator(np.random.PCG64(seed))
    ff = _forcefield(rng, n_species, 12.8, "synthetic-lj")
    box = Box(a, a, a, 90 * DEG, 90 * DEG, 90 * DEG)
    xtal = _framework(rng, f"lj_grid_{seed}", box, n_atoms, n_species, charged=False)
    return Case("lj_grid", xtal, ff, {"CH4": lj_sphere("CH4")},
                meta={"seed": seed, "n_pts": (n_pts, n_pts, n_pts)})


def co2_widom_case(seed: int = 12, n_atoms: int = 600, n_species: int = 5,
                   n_per_block: int = 200000) -> Case:
    rng = np.random.Generator(np.random.PCG64(seed))
    ff = _forcefield(rng, n_species, 12.8, "synthetic-co2")
    lengths = rng.uniform(18.0, 22.0, 3)
    angles = rng.uniform(80.0, 100.0, 3) * DEG
    box = Box(*lengths, *angles)
    xtal = _framework(rng, f"co2_widom_{seed}", box, n_atoms, n_species, charged=True)
    return Case("co2_widom", xtal, ff, {"CO2": co2_molecule()},
                meta={"seed": seed, "n_blocks": 5, "n_per_block": n_per_block,
                      "insertion_seed": 1234567})


def screening_cases(seed: int = 13, n_frameworks: int = 256, density_cap: float = 0.05):
    rng = np.random.Generator(np.random.PCG64(seed))
    ff = _forcefield(rng, 5, 12.8, "synthetic-screening")
    out = []
    for i in range(n_frameworks):
        n = int(round(math.exp(rng.uniform(math.log(50.0), math.log(3000.0)))))
        lengths = rng.uniform(8.0, 40.0, 3)
        angles = rng.uniform(70.0, 110.0, 3) * DEG
        box = Box(*lengths, *angles)
        n_eff = max(1, min(n, int(density_cap * box.Omega)))
        xtal = _framework(rng, f"screen_{seed}_{i}", box, n_eff, 5, charged=True)
        out.append(Case("screening", xtal, ff, {"CH4": lj_sphere("CH4"), "CO2": co2_molecule()},
                        meta={"seed": seed, "index": i, "n_drawn": n}))
    return out


def large_grid_case(seed: int = 14, n_atoms: int = 3000, n_pts: int = 200) -> Case:
    rng = np.random.Generator(np.random.PCG64(seed))
    ff = _forcefield(rng, 5, 12.8, "synthetic-large-grid")
    lengths = rng.uniform(30.0, 40.0, 3)
    angles = rng.uniform(75.0, 105.0, 3) * DEG
    box = Box(*lengths, *angles)
    xtal = _framework(rng, f"large_grid_{seed}", box, n_atoms, 5, charged=True)
    return Case("large_grid", xtal, ff, {"CH4": lj_sphere("CH4"), "ion": point_charge(1.0)},
                meta={"seed": seed, "n_pts": (n_pts, n_pts, n_pts)})


def small_grid_cases(seed: int = 15, n_frameworks: int = 32, n_pts: int = 80):
    rng = np.random.Generator(np.random.PCG64(seed))
    ff = _forcefield(rng, 5, 12.8, "synthetic-small-grids")
    out = []
    for i in range(n_frameworks):
        n = int(rng.integers(200, 1501))
        base = (n / 0.05) ** (1.0 / 3.0)
        lengths = base * rng.uniform(0.9, 1.1, 3)
        angles = rng.uniform(75.0, 105.0, 3) * DEG
        box = Box(*lengths, *angles)
        xtal = _framework(rng, f"small_grid_{seed}_{i}", box, n, 5, charged=True)
        out.append(Case("small_grid", xtal, ff, {"CH4": lj_sphere("CH4"), "ion": point_charge(1.0)},
                        meta={"seed": seed, "index": i, "n_pts": (n_pts, n_pts, n_pts)}))
    return out
