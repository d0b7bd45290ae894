"""ctypes binding of libpmhip.so (declared in include/pmhip.h).

The library is the only compute path: if it is missing or cannot be loaded
every energy call raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
# PMHIP_LIB names another build of the library to load (a sanitizer build of
# its host code, say); by default the in-tree build
LIB_PATH = os.environ.get("PMHIP_LIB") or os.path.join(HERE, "libpmhip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pmhip.h")

PM_OK, PM_ERR_ARG, PM_ERR_DEVICE, PM_ERR_OOM, PM_ERR_STATE = 0, 1, 2, 3, 4
PM_FP64, PM_MIXED = 0, 1
PM_MAX_SITES = 8
PM_MAX_GRAD_TYPES = 8
PM_TUNE_CELL_POINTS, PM_TUNE_CHUNK_R_MAX = 0, 1

_ERR_NAMES = {1: "PM_ERR_ARG", 2: "PM_ERR_DEVICE", 3: "PM_ERR_OOM", 4: "PM_ERR_STATE"}


class PMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None

P = ctypes.c_void_p
PP = ctypes.POINTER(ctypes.c_void_p)
D = ctypes.POINTER(ctypes.c_double)
I32 = ctypes.POINTER(ctypes.c_int32)
I64 = ctypes.POINTER(ctypes.c_int64)
U8 = ctypes.POINTER(ctypes.c_uint8)



class WidomRun(ctypes.Structure):
    """pm_widom_run: one Widom run of pm_widom_batch."""
    _fields_ = [("framework", P), ("forcefield", P), ("ewald", P), ("probe", P),
                ("temperature", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("n_per_block", ctypes.c_int64), ("n_blocks", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("access", P)]


PM_MAX_GCMC_SPECIES, PM_GCMC_BLOCKS, PM_GCMC_MOVES = 4, 5, 6


class GcmcSettings(ctypes.Structure):
    """pm_gcmc_settings."""
    _fields_ = [("n_burn_cycles", ctypes.c_int64), ("n_sample_cycles", ctypes.c_int64),
                ("sample_frequency", ctypes.c_int32), ("snapshot_frequency", ctypes.c_int32),
                ("capacity", ctypes.c_int32), ("density_species", ctypes.c_int32),
                ("density_site_mask", ctypes.c_uint32), ("density_n_pts", ctypes.c_int32 * 3),
                ("density_sim_box", ctypes.c_int32), ("delta", ctypes.c_double), ("seed", ctypes.c_uint64)]


class GcmcState(ctypes.Structure):
    """pm_gcmc_state."""
    _fields_ = [("cycles_done", ctypes.c_int64), ("markov_time", ctypes.c_int64), ("num_snapshots", ctypes.c_int64),
                ("n", ctypes.c_int32 * PM_MAX_GCMC_SPECIES), ("block", ctypes.c_int32), ("error", ctypes.c_int32),
                ("delta", ctypes.c_double), ("energy_gh", ctypes.c_double), ("energy_gg", ctypes.c_double),
                ("n_proposed", ctypes.c_int64 * PM_GCMC_MOVES), ("n_accepted", ctypes.c_int64 * PM_GCMC_MOVES),
                ("block_samples", ctypes.c_int64 * PM_GCMC_BLOCKS),
                ("block_n", (ctypes.c_int64 * PM_MAX_GCMC_SPECIES) * PM_GCMC_BLOCKS),
                ("block_n2", (ctypes.c_int64 * PM_MAX_GCMC_SPECIES) * PM_GCMC_BLOCKS),
                ("block_gh", ctypes.c_double * PM_GCMC_BLOCKS), ("block_gg", ctypes.c_double * PM_GCMC_BLOCKS),
                ("block_gh2", ctypes.c_double * PM_GCMC_BLOCKS), ("block_gg2", ctypes.c_double * PM_GCMC_BLOCKS),
                ("block_un", ctypes.c_double * PM_GCMC_BLOCKS)]


class GridRun(ctypes.Structure):
    """pm_grid_run: one grid of pm_energy_grid_batch."""
    _fields_ = [("framework", P), ("forcefield", P), ("ewald", P), ("probe", P),
                ("n_pts", ctypes.c_int32 * 3), ("reserved", ctypes.c_int32), ("out", D)]


_SIGS = {
    "pm_version": (ctypes.c_char_p, []),
    "pm_last_error": (ctypes.c_char_p, []),
    "pm_init": (ctypes.c_int, [ctypes.c_int32]),
    "pm_framework_create": (ctypes.c_int, [D, I32, ctypes.c_int64, D, I32, D, PP]),
    "pm_framework_destroy": (ctypes.c_int, [P]),
    "pm_framework_info": (ctypes.c_int, [P, I64, I32, D]),
    "pm_forcefield_create": (ctypes.c_int, [ctypes.c_int32, D, D, ctypes.c_double, PP]),
    "pm_forcefield_destroy": (ctypes.c_int, [P]),
    "pm_ewald_create": (ctypes.c_int, [P, ctypes.c_double, ctypes.c_double, ctypes.c_int64, I32, D, PP]),
    "pm_ewald_destroy": (ctypes.c_int, [P]),
    "pm_ewald_info": (ctypes.c_int, [P, I64]),
    "pm_probe_create": (ctypes.c_int, [ctypes.c_int32, D, I32, ctypes.c_int32, D, D, PP]),
    "pm_probe_destroy": (ctypes.c_int, [P]),
    "pm_access_create": (ctypes.c_int, [I32, U8, PP]),
    "pm_access_destroy": (ctypes.c_int, [P]),
    "pm_block_pockets": (ctypes.c_int, [I32, D, ctypes.c_double, ctypes.c_int32, ctypes.c_int64, U8, I32, I64]),
    "pm_vdw_energy": (ctypes.c_int, [P, P, ctypes.c_int32, I32, ctypes.c_int64, D, ctypes.c_int32, D]),
    "pm_electrostatic_energy": (ctypes.c_int, [P, P, ctypes.c_int32, D, ctypes.c_int64, D,
                                               ctypes.c_int32, D, D]),
    "pm_pose_gradient": (ctypes.c_int, [P, P, P, ctypes.c_int32, I32, D, ctypes.c_int64, D, ctypes.c_int32, D, D]),
    "pm_widom": (ctypes.c_int, [P, P, P, P, ctypes.c_double, ctypes.c_uint64, ctypes.c_int32,
                                ctypes.c_int64, ctypes.c_int32, D, D, P]),
    "pm_widom_noalloc": (ctypes.c_int, [P, P, P, P, ctypes.c_double, ctypes.c_uint64, ctypes.c_int32,
                                        ctypes.c_int64, ctypes.c_int32, D, P]),
    "pm_energy_grid": (ctypes.c_int, [P, P, P, P, I32, ctypes.c_int32, D]),
    "pm_widom_gradient": (ctypes.c_int, [P, P, P, P, ctypes.c_double, ctypes.c_uint64, ctypes.c_int32,
                                         ctypes.c_int64, ctypes.c_int32, D, D, D, P]),
    "pm_widom_batch": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(WidomRun), ctypes.c_int32, D]),
    "pm_energy_grid_batch": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(GridRun), ctypes.c_int32]),
    "pm_set_tuning": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double]),
    "pm_get_tuning": (ctypes.c_int, [ctypes.c_int32, D]),
    "pm_autotune_widom": (ctypes.c_int, [P, P, P, P, ctypes.c_double, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int64,
                                         ctypes.c_int32, D, D, D]),
    "pm_copy_probe": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, D]),
    "pm_gcmc_create": (ctypes.c_int, [P, P, ctypes.c_int32, ctypes.POINTER(P), ctypes.c_int32, D, D,
                                      ctypes.POINTER(GcmcSettings), PP]),
    "pm_gcmc_destroy": (ctypes.c_int, [P]),
    "pm_gcmc_set_molecules": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, D, D]),
    "pm_gcmc_run": (ctypes.c_int, [P, ctypes.c_int64]),
    "pm_gcmc_get_state": (ctypes.c_int, [P, ctypes.c_int32, ctypes.POINTER(GcmcState)]),
    "pm_gcmc_get_molecules": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, I32, D, D]),
    "pm_gcmc_total_energy": (ctypes.c_int, [P, ctypes.c_int32, D, D]),
    "pm_gcmc_get_density": (ctypes.c_int, [P, ctypes.c_int32, D]),
    "pm_gcmc_set_lds_molecules": (ctypes.c_int, [P, ctypes.c_int32]),
    "pm_last_kernel_stats": (ctypes.c_int, [D, D]),
    "pm_last_pair_kernel_stats": (ctypes.c_int, [D, D]),
}


def header_functions(path: str = HEADER):
    """Names of the functions include/pmhip.h declares."""
    with open(path) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z_0-9]+)\s*\(", text)))


def load(path: str = LIB_PATH):
    """Load the library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "from the repository root")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != PM_OK:
        raise PMError(rc, load().pm_last_error().decode(errors="replace"))


def dptr(a):
    return a.ctypes.data_as(D)


def iptr(a):
    return a.ctypes.data_as(I32)
