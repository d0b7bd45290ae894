"""Device handles behind the reference call surface.

Frameworks, force fields, Ewald settings and probes are uploaded once and
cached on the Python objects they come from (the Julia shim does the same with
finalizers).  A crystal produced by ``replicate`` is uploaded as its unit cell
plus replication factors, so the library can fold the replicated box's
k-vectors onto the unit-cell structure factor.
"""
from __future__ import annotations

import ctypes
import math
import threading
import weakref

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr

_init_lock = threading.Lock()
_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    with _init_lock:
        if not _initialised:
            check(_lib.load().pm_init(int(device)))
            _initialised = True


def precision_code(precision) -> int:
    p = str(precision).lower()
    if p in ("mixed", "fp32", "pm_mixed"):
        return _lib.PM_MIXED
    if p in ("fp64", "float64", "double", "pm_fp64"):
        return _lib.PM_FP64
    raise ValueError(f"precision must be 'mixed' or 'fp64', got {precision!r}")


class _Handle:
    def __init__(self, ptr, destroy_name):
        self.ptr = ptr
        self._fin = weakref.finalize(self, _destroy, destroy_name, ptr.value)

    @property
    def value(self):
        return self.ptr


def _destroy(name, raw):
    try:
        getattr(_lib.load(), name)(ctypes.c_void_p(raw))
    except Exception:
        pass


def framework(crystal, ljff=None) -> _Handle:
    """Upload ``crystal`` (species indexed by ``ljff``; all 0 without one)."""
    key = ("fw", tuple(ljff.species) if ljff is not None else None)
    h = crystal._device_cache.get(key)
    if h is not None:
        return h
    init()
    unit = crystal._unit if crystal._unit is not None else crystal
    rep = crystal._rep if crystal._unit is not None else (1, 1, 1)
    if not unit.box.is_upper_triangular:
        raise ValueError("the device library needs an upper-triangular f_to_c (a along x)")
    n = unit.n_atoms
    if ljff is not None:
        sp = np.array([ljff.index(s) for s in unit.species], dtype=np.int32)
    else:
        sp = np.zeros(n, dtype=np.int32)
    xf = np.ascontiguousarray(unit.xf.T, dtype=np.float64)
    q = np.ascontiguousarray(unit.atom_q, dtype=np.float64)
    f = np.ascontiguousarray(unit.box.f_to_c.T.reshape(-1))  # column-major
    r = np.array(rep, dtype=np.int32)
    out = ctypes.c_void_p()
    check(_lib.load().pm_framework_create(dptr(f), iptr(r), n, dptr(xf), iptr(sp), dptr(q),
                                          ctypes.byref(out)))
    h = _Handle(out, "pm_framework_destroy")
    crystal._device_cache[key] = h
    return h


_ff_cache = weakref.WeakKeyDictionary()


def forcefield(ljff) -> _Handle:
    h = _ff_cache.get(ljff)
    if h is not None:
        return h
    init()
    eps = np.ascontiguousarray(ljff.epsilon_table, dtype=np.float64)
    sig2 = np.ascontiguousarray(ljff.sigma2_table, dtype=np.float64)
    out = ctypes.c_void_p()
    check(_lib.load().pm_forcefield_create(len(ljff.species), dptr(eps), dptr(sig2),
                                           float(ljff.r2_cutoff), ctypes.byref(out)))
    h = _Handle(out, "pm_forcefield_destroy")
    _ff_cache[ljff] = h
    return h


def ewald(crystal, fw: _Handle, eparams) -> _Handle:
    kidx = np.ascontiguousarray(eparams.kidx, dtype=np.int32)
    wt = np.ascontiguousarray(eparams.wt, dtype=np.float64)
    # keyed by content (an id() of a collected EwaldParams can be reused)
    key = ("ew", float(eparams.alpha), float(eparams.sr_cutoff_r), kidx.tobytes(), wt.tobytes(), fw.ptr.value)
    h = crystal._device_cache.get(key)
    if h is not None:
        return h
    out = ctypes.c_void_p()
    check(_lib.load().pm_ewald_create(fw.ptr, float(eparams.alpha), float(eparams.sr_cutoff_r),
                                      int(wt.size), iptr(kidx), dptr(wt), ctypes.byref(out)))
    h = _Handle(out, "pm_ewald_destroy")
    h._eparams_ref = eparams  # keep id() stable while cached
    crystal._device_cache[key] = h
    return h


def probe(molecule, ljff) -> _Handle:
    """Rigid template from a Cartesian molecule (offsets from its COM)."""
    from .molecule import site_offsets
    init()
    lj_off, q_off = site_offsets(molecule)
    n_lj, n_q = lj_off.shape[1], q_off.shape[1]
    sp = np.array([ljff.index(s) for s in molecule.atom_species] if n_lj else [0], dtype=np.int32)
    lj = np.ascontiguousarray(lj_off.T if n_lj else np.zeros((1, 3)), dtype=np.float64)
    qx = np.ascontiguousarray(q_off.T if n_q else np.zeros((1, 3)), dtype=np.float64)
    qq = np.ascontiguousarray(molecule.charges_q if n_q else np.zeros(1), dtype=np.float64)
    out = ctypes.c_void_p()
    check(_lib.load().pm_probe_create(n_lj, dptr(lj), iptr(sp), n_q, dptr(qx), dptr(qq),
                                      ctypes.byref(out)))
    return _Handle(out, "pm_probe_destroy")


def access(acc_grid) -> _Handle:
    """Upload an accessibility grid (``Grid`` with boolean ``data`` of shape
    ``n_pts``, i fastest) for Widom runs; cached on the grid object."""
    h = getattr(acc_grid, "_device_handle", None)
    data = np.asarray(acc_grid.data)
    n = np.array(data.shape, dtype=np.int32)
    flat = np.ascontiguousarray(data.reshape(-1, order="F").astype(np.uint8))
    key = (tuple(n), flat.tobytes())
    if h is not None and h[0] == key:
        return h[1]
    init()
    out = ctypes.c_void_p()
    check(_lib.load().pm_access_create(iptr(n), flat.ctypes.data_as(_lib.U8), ctypes.byref(out)))
    handle = _Handle(out, "pm_access_destroy")
    try:
        acc_grid._device_handle = (key, handle)
    except AttributeError:
        pass
    return handle


def block_pockets(data, energy_tol, block_pockets=True, max_cycles=0):
    """Host pocket blocking of an energy grid (pm_block_pockets).

    ``data``: (n1, n2, n3) energies in the units of ``energy_tol``.  Returns
    (accessible bool (n1, n2, n3), segment labels int32 (n1, n2, n3) or None,
    (n_segments, n_blocked, n_cycles))."""
    data = np.asarray(data, dtype=np.float64)
    n = np.array(data.shape, dtype=np.int32)
    flat = np.ascontiguousarray(data.reshape(-1, order="F"))
    acc = np.zeros(flat.size, dtype=np.uint8)
    seg = np.zeros(flat.size, dtype=np.int32) if block_pockets else None
    counts = np.zeros(3, dtype=np.int64)
    check(_lib.load().pm_block_pockets(iptr(n), dptr(flat), float(energy_tol), 1 if block_pockets else 0,
                                       int(max_cycles), acc.ctypes.data_as(_lib.U8),
                                       iptr(seg) if seg is not None else None,
                                       counts.ctypes.data_as(_lib.I64)))
    shape = tuple(int(x) for x in n)
    return (acc.reshape(shape, order="F").astype(bool),
            seg.reshape(shape, order="F") if seg is not None else None,
            tuple(int(c) for c in counts))


def last_kernel_stats():
    ms = ctypes.c_double()
    pairs = ctypes.c_double()
    check(_lib.load().pm_last_kernel_stats(ctypes.byref(ms), ctypes.byref(pairs)))
    return ms.value, pairs.value


def last_pair_kernel_stats():
    """(pair-kernel ms, evaluated candidate pairs) of the last Widom/grid call."""
    ms = ctypes.c_double()
    pairs = ctypes.c_double()
    check(_lib.load().pm_last_pair_kernel_stats(ctypes.byref(ms), ctypes.byref(pairs)))
    return ms.value, pairs.value


def fw_info(fw: _Handle):
    n = ctypes.c_int64()
    ns = ctypes.c_int32()
    q = ctypes.c_double()
    check(_lib.load().pm_framework_info(fw.ptr, ctypes.byref(n), ctypes.byref(ns), ctypes.byref(q)))
    return n.value, ns.value, q.value


def ewald_info(ew: _Handle) -> int:
    n = ctypes.c_int64()
    check(_lib.load().pm_ewald_info(ew.ptr, ctypes.byref(n)))
    return n.value


SITE_CHUNK = 6  # site counts with a kernel instantiation: 1..6, 8


def _chunks(n):
    for s in range(0, n, SITE_CHUNK):
        yield s, min(n, s + SITE_CHUNK)


def vdw_energies(fw, ff, site_species, site_xf, precision) -> np.ndarray:
    """site_xf: (n_poses, n_sites, 3) replicated-box fractional coordinates."""
    ff = _ff_handle(ff)
    site_xf = np.asarray(site_xf, dtype=np.float64)
    n_poses, n_sites = site_xf.shape[0], site_xf.shape[1]
    total_e = np.zeros(n_poses)
    lib = _lib.load()
    for a, b in _chunks(n_sites):
        x = np.ascontiguousarray(site_xf[:, a:b, :])
        sp = np.ascontiguousarray(np.asarray(site_species[a:b], dtype=np.int32))
        out = np.zeros(n_poses)
        check(lib.pm_vdw_energy(fw.ptr, ff.ptr, b - a, iptr(sp), n_poses, dptr(x),
                                precision_code(precision), dptr(out)))
        total_e += out
    return total_e


def es_energies(fw, ew, charges, charge_xf, precision):
    charge_xf = np.asarray(charge_xf, dtype=np.float64)
    n_poses, n_q = charge_xf.shape[0], charge_xf.shape[1]
    sr = np.zeros(n_poses)
    lr = np.zeros(n_poses)
    lib = _lib.load()
    for a, b in _chunks(n_q):
        x = np.ascontiguousarray(charge_xf[:, a:b, :])
        q = np.ascontiguousarray(np.asarray(charges[a:b], dtype=np.float64))
        o_sr = np.zeros(n_poses)
        o_lr = np.zeros(n_poses)
        check(lib.pm_electrostatic_energy(fw.ptr, ew.ptr, b - a, dptr(q), n_poses, dptr(x),
                                          precision_code(precision), dptr(o_sr), dptr(o_lr)))
        sr += o_sr
        lr += o_lr
    return sr, lr


def _ff_handle(ff):
    return ff if ff is None or isinstance(ff, _Handle) else forcefield(ff)


def widom(fw, ff, ew, pr, temperature, seed, n_blocks, n_per_block, precision,
          return_energies=False, acc=None):
    """Block sums (n_blocks, 3) of a Widom run; ``ff`` is a force-field
    handle or an LJForceField; ``acc`` an accessibility-grid handle
    (:func:`access`) or None."""
    lib = _lib.load()
    ff = _ff_handle(ff)
    sums = np.zeros(3 * n_blocks)
    energies = np.zeros(2 * n_blocks * n_per_block) if return_energies else None
    check(lib.pm_widom(fw.ptr, ff.ptr, ew.ptr if ew is not None else None, pr.ptr,
                       float(temperature), int(seed) & 0xFFFFFFFFFFFFFFFF, int(n_blocks),
                       int(n_per_block), precision_code(precision), dptr(sums),
                       dptr(energies) if energies is not None else None,
                       acc.ptr if acc is not None else None))
    sums = sums.reshape(n_blocks, 3)
    if return_energies:
        return sums, energies.reshape(n_blocks * n_per_block, 2)
    return sums


def widom_noalloc(fw, ff, ew, pr, temperature, seed, n_blocks, n_per_block, precision, acc=None):
    """Block sums (n_blocks, 3) of a Widom run through pm_widom_noalloc: the
    same run as :func:`widom` in the library's fixed scratch array, with no
    device allocation (the framework's image list must exist: one
    :func:`widom` call with the same framework and force field builds it)."""
    lib = _lib.load()
    ff = _ff_handle(ff)
    sums = np.zeros(3 * n_blocks)
    check(lib.pm_widom_noalloc(fw.ptr, ff.ptr, ew.ptr if ew is not None else None, pr.ptr,
                               float(temperature), int(seed) & 0xFFFFFFFFFFFFFFFF, int(n_blocks),
                               int(n_per_block), precision_code(precision), dptr(sums),
                               acc.ptr if acc is not None else None))
    return sums.reshape(n_blocks, 3)


def _grid_out(out, n):
    """The caller's output buffer (C-contiguous float64 of n values, written
    in place: a buffer reused across calls skips the operating system's
    zeroing of fresh pages), or a new one."""
    if out is None:
        return np.zeros(n)
    if not (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous
            and out.flags.writeable and out.size == n):
        raise ValueError(f"out must be a writeable C-contiguous float64 array of {n} values")
    return out if out.ndim == 1 else out.reshape(-1)


def energy_grid(fw, ff, ew, pr, n_pts, precision, out=None) -> np.ndarray:
    lib = _lib.load()
    ff = _ff_handle(ff)
    n = np.array(n_pts, dtype=np.int32)
    out = _grid_out(out, int(np.prod(n_pts)))
    check(lib.pm_energy_grid(fw.ptr, ff.ptr if ff is not None else None,
                             ew.ptr if ew is not None else None, pr.ptr, iptr(n),
                             precision_code(precision), dptr(out)))
    return out


def widom_batch(runs, precision):
    """Block sums of many Widom runs in one call (pm_widom_batch).

    ``runs``: sequence of (fw, ff, ew, pr, temperature, seed, n_blocks,
    n_per_block[, acc]), handles as for :func:`widom`.  Returns a list of
    (n_blocks, 3) arrays in the order of ``runs``."""
    lib = _lib.load()
    arr = (_lib.WidomRun * len(runs))()
    keep = []
    total = 0
    for i, run in enumerate(runs):
        fw, ff, ew, pr, temperature, seed, n_blocks, n_per_block = run[:8]
        acc = run[8] if len(run) > 8 else None
        ff = _ff_handle(ff)
        keep.append(ff)
        r = arr[i]
        r.framework = fw.ptr
        r.forcefield = ff.ptr
        r.ewald = ew.ptr if ew is not None else None
        r.probe = pr.ptr
        r.temperature = float(temperature)
        r.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        r.n_per_block = int(n_per_block)
        r.n_blocks = int(n_blocks)
        r.access = acc.ptr if acc is not None else None
        total += int(n_blocks)
    sums = np.zeros(3 * total)
    check(lib.pm_widom_batch(len(runs), arr, precision_code(precision), dptr(sums)))
    out, k = [], 0
    for run in runs:
        nb = int(run[6])
        out.append(sums[3 * k:3 * (k + nb)].reshape(nb, 3))
        k += nb
    return out


def energy_grid_batch(runs, precision, outs=None):
    """Many energy grids in one call (pm_energy_grid_batch).

    ``runs``: sequence of (fw, ff, ew, pr, n_pts).  Returns a list of flat
    i-fastest arrays in the order of ``runs`` (``outs``: the caller's buffers,
    one per run, written in place)."""
    lib = _lib.load()
    arr = (_lib.GridRun * len(runs))()
    if outs is not None and len(outs) != len(runs):
        raise ValueError("outs must hold one buffer per run")
    res, keep = [], []
    for i, (fw, ff, ew, pr, n_pts) in enumerate(runs):
        ff = _ff_handle(ff)
        keep.append(ff)
        out = _grid_out(None if outs is None else outs[i], int(np.prod(n_pts)))
        res.append(out)
        r = arr[i]
        r.framework = fw.ptr
        r.forcefield = ff.ptr if ff is not None else None
        r.ewald = ew.ptr if ew is not None else None
        r.probe = pr.ptr
        for d in range(3):
            r.n_pts[d] = int(n_pts[d])
        r.out = dptr(out)
    check(lib.pm_energy_grid_batch(len(runs), arr, precision_code(precision)))
    return res


def widom_gradient(fw, ff, ew, pr, temperature, seed, n_blocks, n_per_block, precision, ljff, acc=None):
    """Block sums (n_blocks, 3) and Boltzmann-weighted parameter-derivative
    sums (n_blocks, n_species, 3) of a Widom run (pm_widom_gradient):
    grads[b, t] = sum_i w_i dE_i/d(eps_t, sigma_t, q_t) for framework species t
    of ``ljff`` (its mixing rule supplies the Jacobian)."""
    lib = _lib.load()
    ffh = _ff_handle(ff if ff is not None else ljff)
    n = len(ljff.species)
    jac = np.ascontiguousarray(ljff.mixing_jacobian(), dtype=np.float64)
    sums = np.zeros(3 * n_blocks)
    grads = np.zeros(n_blocks * n * 3)
    check(lib.pm_widom_gradient(fw.ptr, ffh.ptr, ew.ptr if ew is not None else None, pr.ptr,
                                float(temperature), int(seed) & 0xFFFFFFFFFFFFFFFF, int(n_blocks),
                                int(n_per_block), precision_code(precision), dptr(jac), dptr(sums),
                                dptr(grads), acc.ptr if acc is not None else None))
    return sums.reshape(n_blocks, 3), grads.reshape(n_blocks, n, 3)


def pose_gradients(fw, ff, ew, site_species, site_charges, site_xf, precision):
    """Energies (n_poses, 3: vdw, sr, lr) and dE/d(site Cartesian xyz)
    (n_poses, n_sites, 3) in K/A of rigid poses (pm_pose_gradient).
    ``site_species``: force-field index per site or -1 (no LJ sphere);
    ``site_charges``: charge per site (0: none); ``site_xf``: (n_poses,
    n_sites, 3) replicated-box fractional coordinates."""
    lib = _lib.load()
    ffh = _ff_handle(ff)
    site_xf = np.asarray(site_xf, dtype=np.float64)
    n_poses, n_sites = site_xf.shape[0], site_xf.shape[1]
    energies = np.zeros((n_poses, 3))
    grads = np.zeros((n_poses, n_sites, 3))
    for a, b in _chunks(n_sites):
        x = np.ascontiguousarray(site_xf[:, a:b, :])
        sp = np.ascontiguousarray(np.asarray(site_species[a:b], dtype=np.int32))
        q = np.ascontiguousarray(np.asarray(site_charges[a:b], dtype=np.float64))
        e = np.zeros(3 * n_poses)
        g = np.zeros(n_poses * (b - a) * 3)
        check(lib.pm_pose_gradient(fw.ptr, ffh.ptr if ffh is not None else None,
                                   ew.ptr if ew is not None else None, b - a, iptr(sp), dptr(q), n_poses,
                                   dptr(x), precision_code(precision), dptr(e), dptr(g)))
        energies += e.reshape(n_poses, 3)
        grads[:, a:b, :] = g.reshape(n_poses, b - a, 3)
    return energies, grads


def set_tuning(cell_points=None, chunk_r_max=None):
    """Process-wide point-engine tuning (no effect on results)."""
    lib = _lib.load()
    init()
    if cell_points is not None:
        check(lib.pm_set_tuning(_lib.PM_TUNE_CELL_POINTS, float(cell_points)))
    if chunk_r_max is not None:
        check(lib.pm_set_tuning(_lib.PM_TUNE_CHUNK_R_MAX, float(chunk_r_max)))


def get_tuning():
    lib = _lib.load()
    v = ctypes.c_double()
    check(lib.pm_get_tuning(_lib.PM_TUNE_CELL_POINTS, ctypes.byref(v)))
    cp = v.value
    check(lib.pm_get_tuning(_lib.PM_TUNE_CHUNK_R_MAX, ctypes.byref(v)))
    return {"cell_points": cp, "chunk_r_max": v.value}


def autotune_widom(fw, ff, ew, pr, temperature, seed, n_blocks, n_per_block, precision):
    """Time the run for a grid of chunk shapes, keep the fastest
    (pm_autotune_widom); returns {cell_points, chunk_r_max, ms}."""
    lib = _lib.load()
    ffh = _ff_handle(ff)
    cp, cr, ms = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib.pm_autotune_widom(fw.ptr, ffh.ptr, ew.ptr if ew is not None else None, pr.ptr, float(temperature),
                                int(seed) & 0xFFFFFFFFFFFFFFFF, int(n_blocks), int(n_per_block),
                                precision_code(precision), ctypes.byref(cp), ctypes.byref(cr), ctypes.byref(ms)))
    return {"cell_points": cp.value, "chunk_r_max": cr.value, "ms": ms.value}


class GcmcChains:
    """Grand-canonical Markov chains on the device (a pm_gcmc handle): one
    replicated framework, ``templates`` (Molecule objects; their LJ spheres
    are the simulated molecule), one (temperature, fugacities in Pa) per
    chain.  ``density``: None or dict(species, site_mask, n_pts, sim_box)."""

    def __init__(self, fw, ljff, templates, temperatures, fugacities, n_burn_cycles, n_sample_cycles,
                 sample_frequency=1, snapshot_frequency=1, capacity=1024, density=None, delta=2.0, seed=0):
        init()
        lib = _lib.load()
        ffh = _ff_handle(ljff)
        self.templates = list(templates)
        self.n_species = len(self.templates)
        self.n_sites = [m.n_atoms for m in self.templates]
        self._probes = [probe(m, ljff) for m in self.templates]
        temps = np.ascontiguousarray(np.asarray(temperatures, dtype=np.float64).reshape(-1))
        self.n_chains = temps.size
        fug = np.ascontiguousarray(np.asarray(fugacities, dtype=np.float64).reshape(self.n_chains, self.n_species))
        st = _lib.GcmcSettings()
        st.n_burn_cycles = int(n_burn_cycles)
        st.n_sample_cycles = int(n_sample_cycles)
        st.sample_frequency = int(sample_frequency)
        st.snapshot_frequency = int(snapshot_frequency)
        st.capacity = int(capacity)
        st.density_species = -1
        if density is not None:
            st.density_species = int(density["species"])
            st.density_site_mask = int(density["site_mask"])
            for d in range(3):
                st.density_n_pts[d] = int(density["n_pts"][d])
            st.density_sim_box = 1 if density.get("sim_box", True) else 0
        st.delta = float(delta)
        st.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.capacity = int(capacity)
        self.density_n_pts = tuple(int(x) for x in density["n_pts"]) if density is not None else None
        arr = (ctypes.c_void_p * self.n_species)(*[p.ptr.value for p in self._probes])
        out = ctypes.c_void_p()
        check(lib.pm_gcmc_create(fw.ptr, ffh.ptr, self.n_species, arr, self.n_chains, dptr(temps), dptr(fug),
                                 ctypes.byref(st), ctypes.byref(out)))
        self._fw, self._ff = fw, ffh  # keep the handles alive
        self._h = _Handle(out, "pm_gcmc_destroy")

    def set_molecules(self, chain, species, site_xf, com_xf):
        site_xf = np.ascontiguousarray(np.asarray(site_xf, dtype=np.float64).reshape(-1, self.n_sites[species], 3))
        com_xf = np.ascontiguousarray(np.asarray(com_xf, dtype=np.float64).reshape(-1, 3))
        n = site_xf.shape[0]
        if com_xf.shape[0] != n:
            raise ValueError("one centre of mass per molecule")
        check(_lib.load().pm_gcmc_set_molecules(self._h.ptr, int(chain), int(species), n, dptr(site_xf),
                                                dptr(com_xf)))

    def run(self, n_cycles):
        check(_lib.load().pm_gcmc_run(self._h.ptr, int(n_cycles)))

    def set_lds_molecules(self, max_per_species):
        """Most molecules per species a launch keeps in LDS (-1: as many as
        fit, 0: none); a schedule setting that never changes a result."""
        check(_lib.load().pm_gcmc_set_lds_molecules(self._h.ptr, int(max_per_species)))

    def state(self, chain):
        st = _lib.GcmcState()
        check(_lib.load().pm_gcmc_get_state(self._h.ptr, int(chain), ctypes.byref(st)))
        out = {}
        for name, _ in st._fields_:
            v = getattr(st, name)
            out[name] = np.ctypeslib.as_array(v).copy() if not isinstance(v, (int, float)) else v
        return out

    def molecules(self, chain, species):
        lib = _lib.load()
        n = ctypes.c_int32()
        check(lib.pm_gcmc_get_molecules(self._h.ptr, int(chain), int(species), 0, ctypes.byref(n), None, None))
        ns = self.n_sites[species]
        sites = np.zeros((max(n.value, 1), ns, 3))
        com = np.zeros((max(n.value, 1), 3))
        check(lib.pm_gcmc_get_molecules(self._h.ptr, int(chain), int(species), n.value, ctypes.byref(n),
                                        dptr(sites), dptr(com)))
        return sites[:n.value], com[:n.value]

    def total_energy(self, chain):
        gh, gg = ctypes.c_double(), ctypes.c_double()
        check(_lib.load().pm_gcmc_total_energy(self._h.ptr, int(chain), ctypes.byref(gh), ctypes.byref(gg)))
        return gh.value, gg.value

    def density(self, chain):
        if self.density_n_pts is None:
            raise ValueError("these chains track no density grid")
        out = np.zeros(int(np.prod(self.density_n_pts)))
        check(_lib.load().pm_gcmc_get_density(self._h.ptr, int(chain), dptr(out)))
        return out.reshape(self.density_n_pts, order="F")
