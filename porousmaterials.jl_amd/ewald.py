"""Ewald summation settings for the guest-host electrostatic energy.

Mirrors ``src/electrostatics.jl``: the unit constant ``eps0`` (``:6-7``),
``EwaldSum`` (``:9-16``), ``Kvector``/``EwaldParams`` (``:45-63``),
``determine_ewald_params`` (``:123-141``), ``required_kreps`` (``:157-175``),
``precompute_kvec_wts`` (``:198-232``) and ``setup_Ewald_sum`` (``:256-272``).
``Eikr`` (``:90-101``) is the reference's per-call scratch for e^{ik.r}; the
device library evaluates the Fourier sum from a structure factor instead, so
``Eikr`` here only records its shape for call-surface compatibility.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .box import Box

# vacuum permittivity in e^2 / (A K), evaluated in the reference's order
EPS0 = 8.854187817e-12 / (1.602176565e-19 ** 2) / 1.0e10 * 1.3806488e-23
FOUR_PI_EPS0 = 4.0 * math.pi * EPS0


@dataclass
class EwaldSum:
    sr: float = 0.0  # real-space (short-range) part, K
    lr: float = 0.0  # Fourier (long-range) part, K

    def __add__(self, o: "EwaldSum") -> "EwaldSum":
        return EwaldSum(self.sr + o.sr, self.lr + o.lr)

    def __rmul__(self, a: float) -> "EwaldSum":
        return EwaldSum(a * self.sr, a * self.lr)

    __mul__ = __rmul__


def total(x) -> float:
    """``total(::EwaldSum)`` = sr + lr."""
    return float(x.sr + x.lr)


@dataclass
class EwaldParams:
    kreps: tuple
    alpha: float
    sr_cutoff_r: float
    kidx: np.ndarray = field(repr=False)  # (n_k, 3) int32, half-space enumeration order
    wt: np.ndarray = field(repr=False)    # (n_k,) weights

    @property
    def n_kvecs(self) -> int:
        return int(self.wt.size)


@dataclass
class Eikr:
    n_charges: int
    kreps: tuple


def make_eikr(n_charges: int, eparams: EwaldParams) -> Eikr:
    return Eikr(int(n_charges), tuple(eparams.kreps))


def _bisect(f, lo: float, hi: float) -> float:
    """Bisection of a sign-changing bracket down to adjacent doubles."""
    flo = f(lo)
    for _ in range(4000):
        mid = lo + 0.5 * (hi - lo)
        if mid <= lo or mid >= hi:
            break
        fm = f(mid)
        if fm == 0.0:
            return mid
        if (fm > 0.0) == (flo > 0.0):
            lo, flo = mid, fm
        else:
            hi = mid
    return hi if abs(f(hi)) < abs(f(lo)) else lo


def determine_ewald_params(sr_cutoff_r: float, eps: float):
    def sr_error(alpha):
        return math.erfc(sr_cutoff_r * alpha) / sr_cutoff_r - eps

    alpha = _bisect(sr_error, 0.0, 25.0)

    def lr_error(k2):
        if k2 == 0.0:
            return math.inf
        return 2.0 * math.exp(-k2 / (4.0 * alpha ** 2)) / k2 - eps

    max_k2 = _bisect(lr_error, 0.0, 100.0)
    return alpha, max_k2


def required_kreps(box: Box, max_mag_k_sqrd: float):
    kreps = [0, 0, 0]
    rl = box.reciprocal_lattice
    for abc in range(3):
        n = np.zeros(3)
        while True:
            n[abc] += 1
            k = rl.T @ n
            if float(np.dot(k, k)) > max_mag_k_sqrd:
                kreps[abc] = int(n[abc]) - 1
                break
    return tuple(kreps)


def precompute_kvec_wts(kreps, box: Box, alpha: float, max_mag_k_sqrd: float = math.inf):
    """Half-space k-vectors in the reference's enumeration order (ka outer,
    kb, kc inner) with weight 2 e^{-|k|^2/4 alpha^2} / |k|^2 / eps0 / Omega."""
    ka = np.arange(0, kreps[0] + 1)
    kb = np.arange(-kreps[1], kreps[1] + 1)
    kc = np.arange(-kreps[2], kreps[2] + 1)
    A, B, C = np.meshgrid(ka, kb, kc, indexing="ij")
    A, B, C = A.ravel(), B.ravel(), C.ravel()
    keep = ~((A == 0) & (B == 0) & (C == 0))
    keep &= ~((A == 0) & (B < 0))
    keep &= ~((A == 0) & (B == 0) & (C < 0))
    n = np.stack([A[keep], B[keep], C[keep]], axis=1).astype(np.float64)
    k = n @ box.reciprocal_lattice  # rows: transpose(reciprocal_lattice) * n
    k2 = np.einsum("ij,ij->i", k, k)
    sel = k2 < max_mag_k_sqrd
    k2 = k2[sel]
    wt = 2.0 * np.exp(-k2 / (4.0 * alpha ** 2)) / k2 / EPS0 / box.Omega
    return n[sel].astype(np.int32), wt


def setup_Ewald_sum(box: Box, sr_cutoff_r: float, eps: float = 1e-6, verbose: bool = False) -> EwaldParams:
    alpha, max_k2 = determine_ewald_params(sr_cutoff_r, eps)
    kreps = required_kreps(box, max_k2)
    kidx, wt = precompute_kvec_wts(kreps, box, alpha, max_k2)
    ep = EwaldParams(kreps, alpha, float(sr_cutoff_r), kidx, wt)
    if verbose:
        print("\tEwald summation parameters:")
        print(f"\t\tk-replication factors: {kreps[0]} {kreps[1]} {kreps[2]}")
        print(f"\t\tEwald convergence param. alpha = {alpha:f}")
        print(f"\t\tshort-range cutoff radius (A): {sr_cutoff_r:f}")
        print(f"\t\t{ep.n_kvecs} kvectors")
        print(f"\t\t(calculated with specified precision {eps:e})")
    return ep
