"""Lennard-Jones force field, mixing rules and replication factors.

Mirrors ``LJForceField`` (``src/forcefield.jl:12-100``): pure per-species
sigma/epsilon from a CSV with columns ``atom,sigma(A),epsilon(K)`` (any column
order, ``#`` comments), cross terms by Lorentz-Berthelot (``:75-77``), Kong
(``:78-89``) or geometric (``:90-92``) mixing, and ``r2_cutoff = r_cutoff^2``.
The dense per-pair tables handed to the device library are built here.
"""
from __future__ import annotations

import csv
import cmath
import math
import os

import numpy as np

from .box import Box


class LJForceField:
    def __init__(self, name, pure_sigma: dict, pure_epsilon: dict, r_cutoff: float = 14.0,
                 mixing_rules: str = "Lorentz-Berthelot"):
        rule = mixing_rules.lower()
        if rule not in ("lorentz-berthelot", "kong", "geometric"):
            raise ValueError(f"{mixing_rules} mixing rules not implemented")
        self.name = str(name)
        self.mixing_rules = mixing_rules
        self.species = list(pure_sigma.keys())  # fixed order = table index
        self.pure_sigma = {k: float(v) for k, v in pure_sigma.items()}
        self.pure_epsilon = {k: float(pure_epsilon[k]) for k in self.species}
        self.r2_cutoff = float(r_cutoff) ** 2
        n = len(self.species)
        eps = np.zeros((n, n))
        sig2 = np.zeros((n, n))
        for i, a in enumerate(self.species):
            for j, b in enumerate(self.species):
                e, s2 = _mix(rule, self.pure_epsilon[a], self.pure_sigma[a], self.pure_epsilon[b],
                             self.pure_sigma[b])
                eps[i, j] = e
                sig2[i, j] = s2
        self.epsilon_table = eps
        self.sigma2_table = sig2
        self._index = {s: k for k, s in enumerate(self.species)}

    @classmethod
    def from_csv(cls, path: str, r_cutoff: float = 14.0, mixing_rules: str = "Lorentz-Berthelot"):
        sig, eps = {}, {}
        with open(path) as f:
            rows = [ln for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
        reader = csv.DictReader(rows)
        for row in reader:
            atom = row["atom"].strip()
            if atom in sig:
                raise ValueError(f"Atom {atom} listed at least twice in {path}")
            sig[atom] = float(row["sigma(A)"])
            eps[atom] = float(row["epsilon(K)"])
        name = os.path.splitext(os.path.basename(path))[0]
        return cls(name, sig, eps, r_cutoff, mixing_rules)

    def mixing_jacobian(self) -> np.ndarray:
        """(n, n, 4) derivatives of the mixed pair parameters with respect to
        the first species' pure parameters, the second species' held fixed:
        d eps_ij/d eps_i, d eps_ij/d sigma_i, d sigma_ij/d eps_i,
        d sigma_ij/d sigma_i (sigma_ij = sqrt(sigma2_ij)).  Complex-step
        differentiation of the mixing formulas (exact to rounding: every rule
        is analytic in the pure parameters)."""
        rule = self.mixing_rules.lower()
        n = len(self.species)
        jac = np.zeros((n, n, 4))
        h = 1e-30
        for i, a in enumerate(self.species):
            ea, sa = self.pure_epsilon[a], self.pure_sigma[a]
            for j, b in enumerate(self.species):
                eb, sb = self.pure_epsilon[b], self.pure_sigma[b]
                e1, s21 = _mix(rule, complex(ea, h), sa, eb, sb)
                e2, s22 = _mix(rule, ea, complex(sa, h), eb, sb)
                jac[i, j, 0] = e1.imag / h
                jac[i, j, 1] = e2.imag / h
                jac[i, j, 2] = cmath.sqrt(s21).imag / h
                jac[i, j, 3] = cmath.sqrt(s22).imag / h
        return jac

    def index(self, species: str) -> int:
        try:
            return self._index[species]
        except KeyError:
            raise KeyError(f"{species} missing from {self.name} force field") from None

    def epsilon(self, a: str, b: str) -> float:
        return float(self.epsilon_table[self.index(a), self.index(b)])

    def sigma2(self, a: str, b: str) -> float:
        return float(self.sigma2_table[self.index(a), self.index(b)])

    def __repr__(self) -> str:
        return (f"LJForceField({self.name!r}, {len(self.species)} species, "
                f"r_cutoff={math.sqrt(self.r2_cutoff):.4f} A, {self.mixing_rules})")


def _mix(rule, ea, sa, eb, sb):
    """(epsilon_ij, sigma2_ij) of src/forcefield.jl:75-92; complex-safe."""
    if rule == "lorentz-berthelot":
        return cmath.sqrt(ea * eb) if isinstance(ea * eb, complex) else math.sqrt(ea * eb), ((sa + sb) / 2.0) ** 2
    if rule == "kong":
        es6_a = ea * sa ** 6
        es6_b = eb * sb ** 6
        es12_a = es6_a * sa ** 6
        es12_b = es6_b * sb ** 6
        root = cmath.sqrt if isinstance(es6_a * es6_b, complex) else math.sqrt
        es6 = root(es6_a * es6_b)
        es12 = ((es12_a ** (1 / 13) + es12_b ** (1 / 13)) / 2) ** 13
        return es6 ** 2 / es12, (es12 / es6) ** (1 / 3)
    root = cmath.sqrt if isinstance(ea * eb, complex) else math.sqrt
    return root(ea * eb), sa * sb


def forcefield_coverage(species, ljff: LJForceField) -> bool:
    return all(s in ljff._index for s in species)


def lennard_jones(r2: float, sigma2: float, epsilon: float) -> float:
    """``src/vdw_energetics.jl:19-22``."""
    ratio = (sigma2 / r2) ** 3
    return 4.0 * epsilon * ratio * (ratio - 1.0)


def replication_factors(box, r_cutoff):
    """Smallest supercell whose centre is >= r_cutoff from every face
    (``src/forcefield.jl:120-162``; accepts a Box or Crystal and a cutoff
    or an LJForceField)."""
    if not isinstance(box, Box):
        box = box.box
    if isinstance(r_cutoff, LJForceField):
        r_cutoff = math.sqrt(r_cutoff.r2_cutoff)
    f = box.f_to_c
    a, b, c = f[:, 0].copy(), f[:, 1].copy(), f[:, 2].copy()
    n_ab, n_ac, n_bc = np.cross(a, b), np.cross(a, c), np.cross(b, c)
    rep = [1, 1, 1]

    def centre():
        return (a * 0.5 + b * 0.5) + c * 0.5

    c0 = centre()
    while abs(np.dot(n_bc, c0)) / np.linalg.norm(n_bc) < r_cutoff:
        rep[0] += 1
        a = a + f[:, 0]
        c0 = centre()
    while abs(np.dot(n_ac, c0)) / np.linalg.norm(n_ac) < r_cutoff:
        rep[1] += 1
        b = b + f[:, 1]
        c0 = centre()
    while abs(np.dot(n_ab, c0)) / np.linalg.norm(n_ab) < r_cutoff:
        rep[2] += 1
        c = c + f[:, 2]
        c0 = centre()
    return tuple(rep)
