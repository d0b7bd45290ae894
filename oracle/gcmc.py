"""ctypes wrapper of the oracle's GCMC chain (oracle/gcmc.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/ to replay a device chain on the
CPU.  Everything is built here from raw inputs (lattice parameters,
unit-cell atoms, force-field tables, templates) with the oracle's own box,
replication and proposal-probability code (src/gcmc.jl:404-419).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import oracle as orc

D = ctypes.POINTER(ctypes.c_double)
I32 = ctypes.POINTER(ctypes.c_int32)
NS, MS, NB, NM = 4, 8, 5, 6


class Cfg(ctypes.Structure):
    _fields_ = [("F", ctypes.c_double * 9), ("c2f", ctypes.c_double * 9), ("r2_cut", ctypes.c_double),
                ("omega", ctypes.c_double), ("rep", ctypes.c_double * 3), ("n_atoms", ctypes.c_int64),
                ("atom_xf", D), ("atom_type", I32), ("n_types", ctypes.c_int32), ("eps", D), ("sig2", D),
                ("n_species", ctypes.c_int32), ("capacity", ctypes.c_int32), ("n_sites", ctypes.c_int32 * NS),
                ("rotatable", ctypes.c_int32 * NS), ("site_type", (ctypes.c_int32 * MS) * NS),
                ("off", ((ctypes.c_double * 3) * MS) * NS), ("cum", ctypes.c_double * NM),
                ("n_burn", ctypes.c_int64), ("n_sample", ctypes.c_int64), ("adjust_every", ctypes.c_int64),
                ("cycles_per_block", ctypes.c_int64), ("sample_frequency", ctypes.c_int32),
                ("snapshot_frequency", ctypes.c_int32), ("dens_species", ctypes.c_int32),
                ("dens_sim_box", ctypes.c_int32), ("dens_mask", ctypes.c_uint32), ("dens_n", ctypes.c_int32 * 3),
                ("temperature", ctypes.c_double), ("fug", ctypes.c_double * NS), ("seed", ctypes.c_uint64),
                ("chain", ctypes.c_uint32)]


class State(ctypes.Structure):
    _fields_ = [("outer", ctypes.c_int64), ("t", ctypes.c_int64), ("n_trans_prop", ctypes.c_int64),
                ("n_trans_acc", ctypes.c_int64), ("num_snapshots", ctypes.c_int64), ("n", ctypes.c_int32 * NS),
                ("block", ctypes.c_int32), ("error", ctypes.c_int32), ("delta", ctypes.c_double),
                ("e_gh", ctypes.c_double), ("e_gg", ctypes.c_double), ("n_prop", ctypes.c_int64 * NM),
                ("n_acc", ctypes.c_int64 * NM), ("blk_samples", ctypes.c_int64 * NB),
                ("blk_n", (ctypes.c_int64 * NS) * NB), ("blk_n2", (ctypes.c_int64 * NS) * NB),
                ("blk_gh", ctypes.c_double * NB), ("blk_gg", ctypes.c_double * NB),
                ("blk_gh2", ctypes.c_double * NB), ("blk_gg2", ctypes.c_double * NB),
                ("blk_un", ctypes.c_double * NB)]


def _lib():
    L = orc.lib()
    if not getattr(L, "_gcmc_sigs", False):
        L.orc_gcmc_run.restype = ctypes.c_int32
        L.orc_gcmc_run.argtypes = [ctypes.POINTER(Cfg), ctypes.c_int64, ctypes.POINTER(State), D, D, D]
        L.orc_gcmc_total_energy.restype = None
        L.orc_gcmc_total_energy.argtypes = [ctypes.POINTER(Cfg), I32, D, D, D]
        L._gcmc_sigs = True
    return L


def proposal_cumulative(n_species, any_rotatable):
    """src/gcmc.jl:404-419 in the move order (insertion, deletion,
    translation, rotation, reinsertion, identity change); entries from the
    last move with nonzero probability on are 2.0."""
    p = [0.35, 0.35, 0.25, 0.25 if any_rotatable else 0.0, 0.05, 0.1 if n_species > 1 else 0.0]
    tot = sum(p)
    p = [v / tot for v in p]
    cum, acc, last = [], 0.0, 0
    for k, v in enumerate(p):
        acc += v
        cum.append(acc)
        if v > 0.0:
            last = k
    for k in range(last, NM):
        cum[k] = 2.0
    return cum


class OracleChain:
    """One GCMC chain.

    lattice: unit cell (a, b, c, alpha, beta, gamma) in A / radians;
    unit_xf (N, 3); species (N,) force-field indices; eps, sig2 (n x n)
    mixed tables; r_cutoff; templates: list of (offsets (n_s, 3) Cartesian
    from the centre of mass, species (n_s,) force-field indices, rotatable).
    """

    def __init__(self, lattice, unit_xf, species, eps, sig2, r_cutoff, templates, temperature, fugacities,
                 n_burn, n_sample, seed, chain=0, capacity=1024, sample_frequency=1, snapshot_frequency=1,
                 delta=2.0, density=None, rep=None):
        a, b, c, al, be, ga = lattice
        F_unit = orc.box_f_to_c(a, b, c, al, be, ga)
        self.rep = tuple(rep) if rep is not None else orc.replication_factors(F_unit, r_cutoff)
        R = self.rep
        self.F = orc.box_f_to_c(a * R[0], b * R[1], c * R[2], al, be, ga)
        unit_xf = np.ascontiguousarray(np.asarray(unit_xf, dtype=np.float64).reshape(-1, 3))
        self.xf = np.ascontiguousarray(orc.replicate(unit_xf, R))
        nr = int(np.prod(R))
        self.types = np.ascontiguousarray(np.tile(np.asarray(species, dtype=np.int32), nr))
        self.eps = np.ascontiguousarray(eps, dtype=np.float64)
        self.sig2 = np.ascontiguousarray(sig2, dtype=np.float64)
        cfg = Cfg()
        Fc = self.F.T.reshape(-1)
        c2f = np.linalg.inv(self.F).T.reshape(-1)
        for i in range(9):
            cfg.F[i] = Fc[i]
            cfg.c2f[i] = c2f[i]
        cfg.r2_cut = float(r_cutoff) ** 2
        cfg.omega = abs(float(np.linalg.det(self.F)))
        for d in range(3):
            cfg.rep[d] = float(R[d])
        cfg.n_atoms = self.xf.shape[0]
        cfg.atom_xf = self.xf.ctypes.data_as(D)
        cfg.atom_type = self.types.ctypes.data_as(I32)
        cfg.n_types = self.eps.shape[0]
        cfg.eps = self.eps.ctypes.data_as(D)
        cfg.sig2 = self.sig2.ctypes.data_as(D)
        S = len(templates)
        self.n_species = S
        cfg.n_species = S
        cfg.capacity = int(capacity)
        self.n_sites = []
        for s, (off, sp, rot) in enumerate(templates):
            off = np.asarray(off, dtype=np.float64).reshape(-1, 3)
            cfg.n_sites[s] = off.shape[0]
            self.n_sites.append(off.shape[0])
            cfg.rotatable[s] = 1 if rot else 0
            for k in range(off.shape[0]):
                cfg.site_type[s][k] = int(sp[k])
                for i in range(3):
                    cfg.off[s][k][i] = off[k, i]
        for k, v in enumerate(proposal_cumulative(S, any(t[2] for t in templates))):
            cfg.cum[k] = v
        cfg.n_burn = int(n_burn)
        cfg.n_sample = int(n_sample)
        cfg.adjust_every = int(n_burn) // 12
        cfg.cycles_per_block = int(n_sample) // NB
        cfg.sample_frequency = int(sample_frequency)
        cfg.snapshot_frequency = int(snapshot_frequency)
        cfg.dens_species = -1
        self.dens_shape = None
        if density is not None:
            cfg.dens_species = int(density["species"])
            cfg.dens_mask = int(density["site_mask"])
            cfg.dens_sim_box = 1 if density.get("sim_box", True) else 0
            for d in range(3):
                cfg.dens_n[d] = int(density["n_pts"][d])
            self.dens_shape = tuple(int(x) for x in density["n_pts"])
        cfg.temperature = float(temperature)
        for s in range(S):
            cfg.fug[s] = float(fugacities[s])
        cfg.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        cfg.chain = int(chain)
        self.cfg = cfg
        self.capacity = int(capacity)
        self.state = State()
        self.state.delta = float(delta)
        self.sites = np.zeros(NS * self.capacity * MS * 3)
        self.com = np.zeros(NS * self.capacity * 3)
        self.dens = np.zeros(int(np.prod(self.dens_shape))) if self.dens_shape else None

    def set_molecules(self, species, site_xf, com_xf):
        site_xf = np.asarray(site_xf, dtype=np.float64).reshape(-1, self.n_sites[species], 3)
        com_xf = np.asarray(com_xf, dtype=np.float64).reshape(-1, 3)
        s3 = self.sites.reshape(NS, self.capacity, MS, 3)
        c3 = self.com.reshape(NS, self.capacity, 3)
        n = site_xf.shape[0]
        s3[species, :n, :self.n_sites[species], :] = site_xf
        c3[species, :n, :] = com_xf
        self.state.n[species] = n
        gh, gg = self.total_energy()
        self.state.e_gh = gh
        self.state.e_gg = gg

    def run(self, n_cycles):
        return _lib().orc_gcmc_run(ctypes.byref(self.cfg), int(n_cycles), ctypes.byref(self.state),
                                   self.sites.ctypes.data_as(D), self.com.ctypes.data_as(D),
                                   self.dens.ctypes.data_as(D) if self.dens is not None else None)

    def total_energy(self):
        n = np.array(list(self.state.n), dtype=np.int32)
        gh, gg = ctypes.c_double(), ctypes.c_double()
        _lib().orc_gcmc_total_energy(ctypes.byref(self.cfg), n.ctypes.data_as(I32), self.sites.ctypes.data_as(D),
                                     ctypes.byref(gh), ctypes.byref(gg))
        return gh.value, gg.value

    def molecules(self, species):
        n = self.state.n[species]
        s3 = self.sites.reshape(NS, self.capacity, MS, 3)
        c3 = self.com.reshape(NS, self.capacity, 3)
        return s3[species, :n, :self.n_sites[species], :].copy(), c3[species, :n, :].copy()

    def density(self):
        return self.dens.reshape(self.dens_shape, order="F").copy()


def kb_pa_a3():
    """Boltzmann constant in Pa A^3 / K (src/PorousMaterials.jl:14)."""
    return 1.38064852e7


def ideal_gas_n(fugacity_pa, omega_a3, temperature):
    """<N> of an ideal gas in a volume: f V / (k_B T) (the closed form the
    reference's ideal-gas GCMC test uses, test/gcmc_long.jl:24-34)."""
    return fugacity_pa * omega_a3 / (kb_pa_a3() * temperature)

