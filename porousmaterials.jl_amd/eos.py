"""Equations of state for fugacities (src/eos.jl): Peng-Robinson and van der
Waals fluids, read from the reference's property tables
(``PengRobinson_fluid_props.csv``, ``VdW_fluid_props.csv``) in
``PATHS["data"]`` (the reference's ``rc[:paths][:data]``).

Peng-Robinson (src/eos.jl:4-71): a = 0.457235 R^2 Tc^2 / Pc,
b = 0.0777961 R Tc / Pc, kappa = 0.37464 + 1.54226 w - 0.26992 w^2,
alpha = (1 + kappa (1 - sqrt(T/Tc)))^2, A = alpha a P / (R T)^2,
B = b P / (R T); Z is the real root of
Z^3 - (1 - B) Z^2 + (A - 2B - 3B^2) Z - (AB - B^2 - B^3) = 0 closest to 1;
ln phi = Z - 1 - ln(Z - B) - A / (sqrt(8) B) ln((Z + (1 + sqrt 2) B) / (Z + (1 - sqrt 2) B)).

van der Waals (src/eos.jl:98-135): the density is the smallest real root of
P rho^3... in the reference's coefficient order (P, -(P b + R T), a, -a b)
for rho; ln f = ln P + (b - a / (R T)) P / (R T).
"""
from __future__ import annotations

import csv
import math
import os
from dataclasses import dataclass

import numpy as np

UNIV_GAS_CONST = 8.3144598e-5  # m^3 bar / (K mol), src/PorousMaterials.jl:12
PATHS = {"data": os.path.join(os.getcwd(), "data")}


def _read_table(name):
    with open(os.path.join(PATHS["data"], name)) as f:
        rows = [r for r in f if r.strip() and not r.lstrip().startswith("#")]
    return list(csv.DictReader(rows))


@dataclass
class PengRobinsonFluid:
    fluid: str
    Tc: float
    Pc: float
    omega: float

    @classmethod
    def from_table(cls, fluid: str) -> "PengRobinsonFluid":
        for r in _read_table("PengRobinson_fluid_props.csv"):
            if r["fluid"] == str(fluid):
                return cls(str(fluid), float(r["Tc(K)"]), float(r["Pc(bar)"]), float(r["acentric_factor"]))
        raise ValueError(f"fluid {fluid} properties not found in {PATHS['data']}PengRobinson_fluid_props.csv")


@dataclass
class VdWFluid:
    fluid: str
    a: float
    b: float

    @classmethod
    def from_table(cls, fluid: str) -> "VdWFluid":
        for r in _read_table("VdW_fluid_props.csv"):
            if r["fluid"] == str(fluid):
                return cls(str(fluid), float(r["a(bar*m^6/mol^2)"]), float(r["b(m^3/mol)"]))
        raise ValueError(f"Fluid {fluid} constants not found in {PATHS['data']}VdW_fluidops.csv")


def _pr_AB(fluid: PengRobinsonFluid, T: float, P: float):
    R = UNIV_GAS_CONST
    a = (0.457235 * R ** 2 * fluid.Tc ** 2) / fluid.Pc
    b = (0.0777961 * R * fluid.Tc) / fluid.Pc
    kappa = 0.37464 + (1.54226 * fluid.omega) - (0.26992 * fluid.omega ** 2)
    alpha = (1 + kappa * (1 - math.sqrt(T / fluid.Tc))) ** 2
    return alpha * a * P / (R ** 2 * T ** 2), b * P / (R * T)


def _real_roots(coeffs_high_first):
    r = np.roots(coeffs_high_first)
    return np.real(r[np.abs(np.imag(r)) <= 1e-12 * np.maximum(1.0, np.abs(r))])


def compressibility_factor(fluid, T: float, P: float) -> float:
    if isinstance(fluid, PengRobinsonFluid):
        A, B = _pr_AB(fluid, T, P)
        z = _real_roots([1.0, -(1.0 - B), A - 2 * B - 3 * B ** 2, -(A * B - B ** 2 - B ** 3)])
        return float(z[np.argmin(np.abs(z - 1.0))])
    R = UNIV_GAS_CONST
    # Polynomial([P, -(P b + R T), a, -a b]) in rho, lowest degree first
    rho = _real_roots([-fluid.a * fluid.b, fluid.a, -(P * fluid.b + R * T), P])
    return float(P / (np.min(rho) * R * T))


def fugacity_coefficient(fluid, T: float, P: float) -> float:
    if isinstance(fluid, PengRobinsonFluid):
        z = compressibility_factor(fluid, T, P)
        A, B = _pr_AB(fluid, T, P)
        log_phi = z - 1.0 - math.log(z - B) + \
            -A / (math.sqrt(8) * B) * math.log((z + (1 + math.sqrt(2)) * B) / (z + (1 - math.sqrt(2)) * B))
        return math.exp(log_phi)
    R = UNIV_GAS_CONST
    log_f = math.log(P) + (fluid.b - fluid.a / (R * T)) * (P / (R * T))
    return math.exp(log_f) / P


def calculate_properties(fluid, T: float, P: float, verbose: bool = True) -> dict:
    """src/eos.jl:176-211: compressibility factor, molar volume (L/mol),
    density (mol/m^3), fugacity (bar), fugacity coefficient."""
    z = compressibility_factor(fluid, T, P)
    rho = P / (z * UNIV_GAS_CONST * T)
    phi = fugacity_coefficient(fluid, T, P)
    props = {"compressibility factor": z, "molar volume (L/mol)": 1000.0 / rho, "density (mol/m³)": rho,
             "fugacity (bar)": phi * P, "fugacity coefficient": phi}
    if verbose:
        print(f"{fluid.fluid} properties at T = {T:f} K, P = {P:f} bar:")
        for k, v in props.items():
            print(f"\t{k}: {v}")
    return props
