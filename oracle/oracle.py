"""ctypes wrapper of liboracle.so, the CPU FP64 restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py as the checker.  It takes raw inputs (lattice
parameters, unit-cell atoms, force-field tables, probe template) and builds
every derived quantity (box, replication, Ewald set-up) with its own code, so
it shares nothing with the product path it checks.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PM_ORACLE_LIB names another build of the oracle (a sanitizer build, say)
LIB_PATH = os.environ.get("PM_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

D = ctypes.POINTER(ctypes.c_double)
I32 = ctypes.POINTER(ctypes.c_int32)
U32 = ctypes.POINTER(ctypes.c_uint32)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_eps0": (ctypes.c_double, []),
            "orc_four_pi_eps0": (ctypes.c_double, []),
            "orc_box_f_to_c": (None, [ctypes.c_double] * 6 + [D]),
            "orc_replication_factors": (None, [D, ctypes.c_double, I32]),
            "orc_replicate": (ctypes.c_int64, [ctypes.c_int64, D, I32, D]),
            "orc_lennard_jones": (ctypes.c_double, [ctypes.c_double] * 3),
            "orc_vdw_energy": (ctypes.c_double, [ctypes.c_int32, D, I32, ctypes.c_int64, D, I32, D,
                                                 ctypes.c_int32, D, D, ctypes.c_double]),
            "orc_total_vdw_gg": (ctypes.c_double, [ctypes.c_int64, D, D, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double]),
            "orc_determine_ewald_params": (None, [ctypes.c_double, ctypes.c_double, D, D]),
            "orc_reciprocal_lattice": (None, [D, D]),
            "orc_required_kreps": (None, [D, ctypes.c_double, I32]),
            "orc_precompute_kvec_wts": (ctypes.c_int64, [I32, D, ctypes.c_double, ctypes.c_double,
                                                         ctypes.c_int64, I32, D]),
            "orc_phi_eikr": (None, [ctypes.c_int64, D, D, D, D, I32, ctypes.c_int64, I32, D,
                                    ctypes.c_double, ctypes.c_double, D, D]),
            "orc_sk_build": (ctypes.c_int64, [ctypes.c_int64, D, D, I32, ctypes.c_int64, I32, D,
                                              I32, D, D]),
            "orc_phi_lr_sk": (ctypes.c_double, [ctypes.c_int64, I32, D, D, I32, D]),
            "orc_phi_sr": (ctypes.c_double, [ctypes.c_int64, D, D, D, D, ctypes.c_double,
                                             ctypes.c_double]),
            "orc_ewald_gg_components": (None, [ctypes.c_int64, D, D, I32, D, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int64, I32, D, D]),
            "orc_philox4x32_10": (None, [U32, U32, U32]),
            "orc_insertion_uniforms": (None, [ctypes.c_uint64, ctypes.c_uint64, D]),
            "orc_arvo_rotation": (None, [D, D]),
            "orc_widom_energies": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64,
                                                    D, D]),
            "orc_widom_block_sums": (None, [ctypes.c_int64, D, D, ctypes.c_double, D]),
            "orc_energy_grid": (ctypes.c_int32, [ctypes.c_void_p, I32, D]),
            "orc_num_threads": (ctypes.c_int32, []),
            "orc_vdw_energy_abs": (ctypes.c_double, [ctypes.c_int32, D, I32, ctypes.c_int64, D, I32, D,
                                                     ctypes.c_int32, D, D, ctypes.c_double]),
            "orc_phi_sr_abs": (ctypes.c_double, [ctypes.c_int64, D, D, D, D, ctypes.c_double,
                                                 ctypes.c_double]),
            "orc_widom_scales": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, D]),
            "orc_widom_param_grad": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                      ctypes.c_int64, D, D, D, D]),
            "orc_pose_grad": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, I32, D, ctypes.c_int64, D, D, D]),
            "orc_energy_grid_scales": (ctypes.c_int32, [ctypes.c_void_p, I32, D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I32)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------ scalars ----
def eps0() -> float:
    return lib().orc_eps0()


def four_pi_eps0() -> float:
    return lib().orc_four_pi_eps0()


def box_f_to_c(a, b, c, alpha, beta, gamma) -> np.ndarray:
    """3x3 f_to_c as a numpy (row, col) matrix."""
    F = np.zeros(9)
    lib().orc_box_f_to_c(a, b, c, alpha, beta, gamma, _d(F))
    return F.reshape(3, 3, order="F")


def _colmajor(M):
    return _c(np.asarray(M, dtype=np.float64).reshape(3, 3).T.reshape(-1))


def replication_factors(F, r_cutoff):
    rep = np.zeros(3, dtype=np.int32)
    lib().orc_replication_factors(_colmajor(F).ctypes.data_as(D), r_cutoff, _i(rep))
    return tuple(int(x) for x in rep)


def replicate(xf_unit, rep):
    """xf_unit: (N, 3) -> (N * prod(rep), 3)"""
    xf = _c(xf_unit).reshape(-1, 3)
    r = np.array(rep, dtype=np.int32)
    out = np.zeros((xf.shape[0] * int(np.prod(rep)), 3))
    lib().orc_replicate(xf.shape[0], _d(xf), _i(r), _d(out))
    return out


def lennard_jones(r2, sig2, eps):
    return lib().orc_lennard_jones(r2, sig2, eps)


def vdw_energy(site_xf, site_species, fw_xf, fw_species, F, eps, sig2, r2_cutoff):
    sx = _c(site_xf).reshape(-1, 3)
    ss = _c(site_species, np.int32)
    fx = _c(fw_xf).reshape(-1, 3)
    fs = _c(fw_species, np.int32)
    e = _c(eps)
    s2 = _c(sig2)
    return lib().orc_vdw_energy(sx.shape[0], _d(sx), _i(ss), fx.shape[0], _d(fx), _i(fs),
                                _d(_colmajor(F)), e.shape[0], _d(e), _d(s2), r2_cutoff)


def total_vdw_gg(xf, F, sig2, eps, r2_cutoff):
    x = _c(xf).reshape(-1, 3)
    return lib().orc_total_vdw_gg(x.shape[0], _d(x), _d(_colmajor(F)), sig2, eps, r2_cutoff)


def determine_ewald_params(sr_cutoff, precision):
    a = ctypes.c_double()
    k = ctypes.c_double()
    lib().orc_determine_ewald_params(sr_cutoff, precision, ctypes.byref(a), ctypes.byref(k))
    return a.value, k.value


def reciprocal_lattice(F):
    R = np.zeros(9)
    lib().orc_reciprocal_lattice(_d(_colmajor(F)), _d(R))
    return R.reshape(3, 3, order="F")


def required_kreps(F, max_k2):
    k = np.zeros(3, dtype=np.int32)
    lib().orc_required_kreps(_d(_colmajor(F)), max_k2, _i(k))
    return tuple(int(x) for x in k)


def precompute_kvec_wts(kreps, F, alpha, max_k2=math.inf):
    kr = np.array(kreps, dtype=np.int32)
    Fc = _colmajor(F)
    n = lib().orc_precompute_kvec_wts(_i(kr), _d(Fc), alpha, max_k2, 0, None, None)
    kidx = np.zeros((n, 3), dtype=np.int32)
    wt = np.zeros(n)
    lib().orc_precompute_kvec_wts(_i(kr), _d(Fc), alpha, max_k2, n, _i(kidx), _d(wt))
    return kidx, wt


def phi_eikr(xq, q, x, F, kreps, kidx, wt, alpha, sr_cutoff):
    xq = _c(xq).reshape(-1, 3)
    qq = _c(q)
    xx = _c(x)
    kr = np.array(kreps, dtype=np.int32)
    ki = _c(kidx, np.int32)
    w = _c(wt)
    sr = ctypes.c_double()
    lr = ctypes.c_double()
    lib().orc_phi_eikr(xq.shape[0], _d(xq), _d(qq), _d(xx), _d(_colmajor(F)), _i(kr), w.size,
                       _i(ki), _d(w), alpha, sr_cutoff, ctypes.byref(sr), ctypes.byref(lr))
    return sr.value, lr.value


def ewald_gg_components(xf, q, mol, F, alpha, sr_cutoff, kidx, wt):
    x = _c(xf).reshape(-1, 3)
    qq = _c(q)
    m = _c(mol, np.int32)
    ki = _c(kidx, np.int32)
    w = _c(wt)
    out = np.zeros(4)
    lib().orc_ewald_gg_components(x.shape[0], _d(x), _d(qq), _i(m), _d(_colmajor(F)), alpha,
                                  sr_cutoff, w.size, _i(ki), _d(w), _d(out))
    return {"real": out[0], "fourier": out[1], "self": out[2], "intra": out[3]}


def philox4x32_10(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data_as(U32), k.ctypes.data_as(U32), o.ctypes.data_as(U32))
    return [int(v) for v in o]


def insertion_uniforms(seed, g):
    u = np.zeros(6)
    lib().orc_insertion_uniforms(seed, g, _d(u))
    return u


def arvo_rotation(u):
    uu = _c(u)
    R = np.zeros(9)
    lib().orc_arvo_rotation(_d(uu), _d(R))
    return R.reshape(3, 3, order="F")


# ---------------------------------------------------- whole systems ----
class _Cfg(ctypes.Structure):
    _fields_ = [
        ("n_rep", ctypes.c_int64), ("xf", D), ("species", I32), ("q", D),
        ("nq", ctypes.c_int64), ("xq", D), ("qq", D),
        ("F", ctypes.c_double * 9), ("c_to_f", ctypes.c_double * 9),
        ("rep", ctypes.c_int32 * 3),
        ("n_species", ctypes.c_int32), ("eps", D), ("sig2", D), ("r2_cutoff", ctypes.c_double),
        ("n_lj", ctypes.c_int32), ("lj_off", D), ("lj_species", I32),
        ("n_q", ctypes.c_int32), ("q_off", D), ("q_site", D),
        ("charged", ctypes.c_int32), ("alpha", ctypes.c_double), ("sr_cutoff", ctypes.c_double),
        ("es_mode", ctypes.c_int32),
        ("nk_sk", ctypes.c_int64), ("m_sk", I32), ("a_re", D), ("a_im", D),
        ("kreps", ctypes.c_int32 * 3), ("nk", ctypes.c_int64), ("kidx", I32), ("wt", D),
        ("seed", ctypes.c_uint64),
    ]


class _ParamDual(ctypes.Structure):
    _fields_ = [("n_ft", ctypes.c_int32), ("ft", I32), ("pure_eps", D), ("pure_sig", D),
                ("n_unit", ctypes.c_int64), ("unit_xf", D), ("unit_q", D), ("unit_species", I32)]


class OracleSystem:
    """A framework + force field + probe, replicated and Ewald-set-up by the
    oracle itself from raw inputs.

    lattice: (a, b, c, alpha, beta, gamma) of the unit cell (radians);
    unit_xf (N, 3); species (N,) indices into eps/sig2 (n x n); q (N,) or None;
    probe: lj_off (n_lj, 3) Cartesian offsets from the COM, lj_species (n_lj,),
    q_off (n_q, 3), q_site (n_q,).
    """

    def __init__(self, lattice, unit_xf, species, q, eps, sig2, r_cutoff, lj_off, lj_species,
                 q_off, q_site, ewald_precision=1e-6, rep=None, ewald_sr_cutoff=None):
        L = lib()
        self._keep = []
        a, b, c, al, be, ga = lattice
        self.F_unit = box_f_to_c(a, b, c, al, be, ga)
        self.rep = tuple(rep) if rep is not None else replication_factors(self.F_unit, r_cutoff)
        R = self.rep
        self.F = box_f_to_c(a * R[0], b * R[1], c * R[2], al, be, ga)
        self.c_to_f = np.linalg.inv(self.F)
        unit_xf = _c(unit_xf).reshape(-1, 3)
        n_unit = unit_xf.shape[0]
        self.unit_xf = unit_xf
        self.unit_species = _c(np.asarray(species, dtype=np.int32), np.int32)
        nr = int(np.prod(R))
        self.xf = replicate(unit_xf, R)
        self.species = _c(np.tile(np.asarray(species, dtype=np.int32), nr), np.int32)
        q_unit = np.zeros(n_unit) if q is None else _c(q)
        self.unit_q = q_unit
        self.q = _c(np.tile(q_unit, nr))
        nz = self.q != 0.0
        self.xq = _c(self.xf[nz])
        self.qq = _c(self.q[nz])
        self.eps = _c(eps)
        self.sig2 = _c(sig2)
        self.r2_cutoff = float(r_cutoff) ** 2
        self.lj_off = _c(lj_off).reshape(-1, 3)
        self.lj_species = _c(lj_species, np.int32).reshape(-1)
        self.q_off = _c(q_off).reshape(-1, 3)
        self.q_site = _c(q_site).reshape(-1)
        self.charged = bool(self.qq.size > 0 and self.q_site.size > 0)
        sr_cut = float(r_cutoff) if ewald_sr_cutoff is None else float(ewald_sr_cutoff)
        self.sr_cutoff = sr_cut
        self.alpha, self.max_k2 = determine_ewald_params(sr_cut, ewald_precision)
        self.kreps = required_kreps(self.F, self.max_k2)
        self.kidx, self.wt = precompute_kvec_wts(self.kreps, self.F, self.alpha, self.max_k2)
        # structure-factor form
        uq_mask = q_unit != 0.0
        uxq = _c(unit_xf[uq_mask])
        uq = _c(q_unit[uq_mask])
        nk = self.wt.size
        self.m_sk = np.zeros((max(nk, 1), 3), dtype=np.int32)
        self.a_re = np.zeros(max(nk, 1))
        self.a_im = np.zeros(max(nk, 1))
        r = np.array(R, dtype=np.int32)
        self.nk_sk = L.orc_sk_build(uq.size, _d(uxq), _d(uq), _i(r), nk, _i(self.kidx), _d(self.wt),
                                    _i(self.m_sk), _d(self.a_re), _d(self.a_im))
        cfg = _Cfg()
        cfg.n_rep = self.xf.shape[0]
        cfg.xf = _d(self.xf)
        cfg.species = _i(self.species)
        cfg.q = _d(self.q)
        cfg.nq = self.qq.size
        cfg.xq = _d(self.xq)
        cfg.qq = _d(self.qq)
        for i, v in enumerate(self.F.T.reshape(-1)):
            cfg.F[i] = v
        for i, v in enumerate(self.c_to_f.T.reshape(-1)):
            cfg.c_to_f[i] = v
        for i in range(3):
            cfg.rep[i] = R[i]
            cfg.kreps[i] = self.kreps[i]
        cfg.n_species = self.eps.shape[0]
        cfg.eps = _d(self.eps)
        cfg.sig2 = _d(self.sig2)
        cfg.r2_cutoff = self.r2_cutoff
        cfg.n_lj = self.lj_off.shape[0]
        cfg.lj_off = _d(self.lj_off)
        cfg.lj_species = _i(self.lj_species)
        cfg.n_q = self.q_site.size
        cfg.q_off = _d(self.q_off)
        cfg.q_site = _d(self.q_site)
        cfg.charged = int(self.charged)
        cfg.alpha = self.alpha
        cfg.sr_cutoff = sr_cut
        cfg.es_mode = 0
        cfg.nk_sk = self.nk_sk
        cfg.m_sk = _i(self.m_sk)
        cfg.a_re = _d(self.a_re)
        cfg.a_im = _d(self.a_im)
        cfg.nk = nk
        cfg.kidx = _i(self.kidx)
        cfg.wt = _d(self.wt)
        self.cfg = cfg

    # -- per point / per pose ---------------------------------------------
    def vdw(self, site_xf, site_species=None):
        sp = self.lj_species if site_species is None else site_species
        return vdw_energy(site_xf, sp, self.xf, self.species, self.F, self.eps, self.sig2,
                          self.r2_cutoff)

    def vdw_scale(self, site_xf, site_species=None):
        """sum of |LJ pair terms| inside the cutoff (condition scale of vdw)."""
        sp = self.lj_species if site_species is None else site_species
        sx = _c(site_xf).reshape(-1, 3)
        ss = _c(sp, np.int32)
        return lib().orc_vdw_energy_abs(sx.shape[0], _d(sx), _i(ss), self.xf.shape[0], _d(self.xf),
                                        _i(self.species), _d(_colmajor(self.F)), self.eps.shape[0],
                                        _d(self.eps), _d(self.sig2), self.r2_cutoff)

    def phi_scale(self, x):
        """sum |q_j erfc(alpha r)/(4 pi eps0 r)| + |lr| per unit probe charge."""
        x = _c(x)
        a = lib().orc_phi_sr_abs(self.qq.size, _d(self.xq), _d(self.qq), _d(x), _d(_colmajor(self.F)),
                                 self.alpha, self.sr_cutoff)
        return a + abs(self.phi(x)[1])

    def phi(self, x, form="sk"):
        """Potential per unit charge at x (replicated fractional): (sr, lr)."""
        x = _c(x)
        if form == "eikr":
            return phi_eikr(self.xq, self.qq, x, self.F, self.kreps, self.kidx, self.wt,
                            self.alpha, self.sr_cutoff)
        L = lib()
        r = np.array(self.rep, dtype=np.int32)
        sr = L.orc_phi_sr(self.qq.size, _d(self.xq), _d(self.qq), _d(x), _d(_colmajor(self.F)),
                          self.alpha, self.sr_cutoff)
        lr = L.orc_phi_lr_sk(self.nk_sk, _i(self.m_sk), _d(self.a_re), _d(self.a_im), _i(r), _d(x))
        return sr, lr

    # -- Widom / grids ------------------------------------------------------
    def widom_energies(self, seed, g0, count, es_mode=0):
        self.cfg.seed = int(seed)
        self.cfg.es_mode = int(es_mode)
        v = np.zeros(count)
        e = np.zeros(count)
        rc = lib().orc_widom_energies(ctypes.byref(self.cfg), int(g0), int(count), _d(v), _d(e))
        if rc != 0:
            raise ValueError("oracle: too many probe sites")
        return v, e

    def widom_scales(self, seed, g0, count):
        """per-insertion condition scale (sum of absolute pair contributions)."""
        self.cfg.seed = int(seed)
        out = np.zeros(count)
        lib().orc_widom_scales(ctypes.byref(self.cfg), int(g0), int(count), _d(out))
        return out

    def energy_grid_scales(self, n_pts):
        n = np.array(n_pts, dtype=np.int32)
        out = np.zeros(int(np.prod(n_pts)))
        lib().orc_energy_grid_scales(ctypes.byref(self.cfg), _i(n), _d(out))
        return out

    def block_sums(self, vdw, es, temperature):
        out = np.zeros(3)
        v = _c(vdw)
        e = _c(es)
        lib().orc_widom_block_sums(v.size, _d(v), _d(e), temperature, _d(out))
        return out

    def energy_grid(self, n_pts):
        n = np.array(n_pts, dtype=np.int32)
        out = np.zeros(int(np.prod(n_pts)))
        lib().orc_energy_grid(ctypes.byref(self.cfg), _i(n), _d(out))
        return out

    # -- dual-number derivatives (oracle/dual.c) -----------------------------
    def widom_param_grad(self, seed, g0, count, pure_eps, pure_sig, ft, cond=False):
        """Insertions [g0, g0 + count) of the stream: (E_vdw, E_es, dE) with
        dE (count, len(ft), 3) = dE/d(eps_t, sigma_t, q_t) for the force-field
        species ft (framework side; Lorentz-Berthelot mixing of the pure
        parameters pure_eps / pure_sig, indexed by force-field species; q_t a
        uniform shift of every framework atom of species t).  cond=True adds a
        fourth array of dE's shape: each derivative's sum of summand
        magnitudes (per pair and per site's Fourier part)."""
        ft = np.asarray(ft, dtype=np.int32)
        pe, ps = _c(pure_eps), _c(pure_sig)
        p = _ParamDual(len(ft), _i(ft), _d(pe), _d(ps), self.unit_xf.shape[0], _d(self.unit_xf),
                       _d(self.unit_q), _i(self.unit_species))
        self.cfg.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        v, e = np.zeros(count), np.zeros(count)
        d = np.zeros((count, len(ft), 3))
        cd = np.zeros((count, len(ft), 3)) if cond else None
        rc = lib().orc_widom_param_grad(ctypes.byref(self.cfg), ctypes.byref(p), int(g0), int(count), _d(v), _d(e),
                                        _d(d), _d(cd) if cond else None)
        if rc:
            raise ValueError("orc_widom_param_grad: too many species or sites")
        return (v, e, d, cd) if cond else (v, e, d)

    def pose_grad(self, site_species, site_q, site_xf):
        """Energies (n_poses,) and dE/d(site Cartesian xyz) (n_poses, n_sites,
        3) of explicit poses, site_xf (n_poses, n_sites, 3) fractional in the
        replicated box; site_species -1 for a site without an LJ sphere."""
        x = _c(site_xf)
        n_poses, n_sites = x.shape[0], x.shape[1]
        sp = _c(np.asarray(site_species, dtype=np.int32), np.int32)
        q = _c(site_q)
        e = np.zeros(n_poses)
        g = np.zeros((n_poses, n_sites, 3))
        if lib().orc_pose_grad(ctypes.byref(self.cfg), n_sites, _i(sp), _d(q), n_poses, _d(x), _d(e), _d(g)):
            raise ValueError("orc_pose_grad: at most 8 sites")
        return e, g
