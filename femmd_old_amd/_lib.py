"""ctypes binding of include/femmd.h (libfemmd.so, built by femmd_old_amd/build.py).

There is no fallback: if the HIP library is missing, importing the package
raises.  Build it first with ``python -m femmd_old_amd.build`` (or
``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfemmd.so"

FM_OK = 0
FM_ERR_ARG = -1
FM_ERR_STATE = -2
FM_ERR_HIP = -3
FM_ERR_DEVICE = -4
FM_ERR_GEOMETRY = -5
FM_ERR_NOMEM = -6

FM_DEV_CELL_OVERFLOW = 1
FM_DEV_TILE_OVERFLOW = 2
FM_DEV_INVERTED = 4
FM_DEV_NONFINITE = 8
FM_DEV_LIST_OVERFLOW = 16
FM_DEV_SORT_STALL = 32

FM_NEO_HOOKEAN = 0
FM_SVK = 1
FM_LEGACY_GPU = 2


class FmParams(C.Structure):
    _fields_ = [
        ("Lx", C.c_double),
        ("Ly", C.c_double),
        ("dt", C.c_double),
        ("device", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class FmEnergies(C.Structure):
    _fields_ = [
        ("ke", C.c_double),
        ("u_lj", C.c_double),
        ("w_lj", C.c_double),
        ("u_bond", C.c_double),
        ("w_bond", C.c_double),
        ("u_fea", C.c_double),
        ("w_fea", C.c_double),
        ("vv_max", C.c_double),
        ("temperature", C.c_double),
        ("pressure", C.c_double),
        ("steps", C.c_int64),
        ("n", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


AVG_NAMES = ("uKinetic", "uPotential", "uTotal", "pressure", "virial", "feaVirial", "uBonded",
             "uFea")


class FmAverages(C.Structure):
    _fields_ = [
        ("count", C.c_int64),
        ("mean", C.c_double * 8),
        ("stdv", C.c_double * 8),
    ]

    def as_dict(self) -> dict:
        d = {"count": self.count}
        for k, name in enumerate(AVG_NAMES):
            d[name] = {"avg": self.mean[k], "stdv": self.stdv[k]}
        return d


_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int32)
_l = C.POINTER(C.c_int64)
_d = C.POINTER(C.c_double)
_h = C.c_void_p

# name -> (restype, argtypes); the full list of entry points of femmd.h
PROTOTYPES = {
    "fm_device_count": (C.c_int, [_i]),
    "fm_create": (C.c_int, [C.POINTER(FmParams), C.POINTER(_h)]),
    "fm_destroy": (None, [_h]),
    "fm_last_error": (C.c_char_p, [_h]),
    "fm_add_particles": (C.c_int, [_h, C.c_int32, _f, _f, _f, _f, _f, _f, _i, _i]),
    "fm_add_elastomer": (C.c_int, [_h, C.c_int32, _f, C.c_int32, _i, C.c_int32, C.c_float,
                                   C.c_float, C.c_int32, C.c_float, C.c_float, C.c_float,
                                   C.c_int32, _f, _f, _i]),
    "fm_add_bonds": (C.c_int, [_h, C.c_int32, _i, _i, _f, _f]),
    "fm_add_edge_bonds": (C.c_int, [_h, C.c_int32, C.c_int32, _i, C.c_int32, C.c_float, C.c_int32]),
    "fm_set_pair_table": (C.c_int, [_h, C.c_int32, _f, _f, _f]),
    "fm_set_material": (C.c_int, [_h, C.c_int32, C.c_float, C.c_float]),
    "fm_set_dt": (C.c_int, [_h, C.c_double]),
    "fm_set_lj": (C.c_int, [_h, C.c_int32]),
    "fm_set_exclusion": (C.c_int, [_h, C.c_int32, C.c_int32]),
    "fm_set_skin": (C.c_int, [_h, C.c_float]),
    "fm_set_reorder_interval": (C.c_int, [_h, C.c_int32]),
    "fm_reorder": (C.c_int, [_h]),
    "fm_reorder_info": (C.c_int, [_h, _i, _l]),
    "fm_list_stats": (C.c_int, [_h, _d, _i, _l]),
    "fm_cluster_count": (C.c_int, [_h, _i]),
    "fm_cluster_props": (C.c_int, [_h, _d, _f, _f]),
    "fm_cluster_virial": (C.c_int, [_h, _d]),
    "fm_set_temperature": (C.c_int, [_h, C.c_double, C.c_int32, _i]),
    "fm_set_temperature_noalloc": (C.c_int, [_h, C.c_double, C.c_int32, _i]),
    "fm_set_mass": (C.c_int, [_h, C.c_int32, _i, _f]),
    "fm_commit": (C.c_int, [_h]),
    "fm_step": (C.c_int, [_h, C.c_int32]),
    "fm_compute_forces": (C.c_int, [_h]),
    "fm_get_energies": (C.c_int, [_h, C.POINTER(FmEnergies)]),
    "fm_set_observe_every": (C.c_int, [_h, C.c_int32]),
    "fm_get_averages": (C.c_int, [_h, C.POINTER(FmAverages)]),
    "fm_reset_averages": (C.c_int, [_h]),
    "fm_get_state": (C.c_int, [_h, _f, _f, _f, _f]),
    "fm_set_state": (C.c_int, [_h, _f, _f, _f, _f]),
    "fm_get_forces": (C.c_int, [_h, _f, _f]),
    "fm_num_particles": (C.c_int, [_h, _i]),
    "fm_num_bonds": (C.c_int, [_h, _i]),
    "fm_get_bonds": (C.c_int, [_h, _i]),
    "fm_num_elements": (C.c_int, [_h, _i]),
    "fm_sync": (C.c_int, [_h]),
    "fm_set_tuning": (C.c_int, [_h, C.c_int32, C.c_int32, C.c_int32]),
    "fm_device_errors": (C.c_uint32, [_h]),
    "fm_clear_errors": (C.c_int, [_h]),
    "fm_get_cells": (C.c_int, [_h, _i, _i, _i, _i, _i]),
    "fm_pair_sets": (C.c_int, [_h, _i, _l, _l]),
    "fm_time_forces": (C.c_int, [_h, C.c_int32, _d]),
    "fm_profile_step": (C.c_int, [_h, C.c_int32, _d]),
    "fm_force_jvp": (C.c_int, [_h, _f, _f, _f, _f, _d, _d, _d]),
    "fm_adjoint": (C.c_int, [_h, C.c_int32, C.c_int32, C.c_int32, _i, _f, _f, _d, _d, _d,
                             _d]),
    "fm_adjoint_forward": (C.c_int, [_h, C.c_int32, C.c_int32]),
    "fm_adjoint_backward": (C.c_int, [_h, _f, _f, _d, _d, _d]),
    "fm_adjoint_run": (C.c_int, [_h, C.c_int32, C.c_int32, _f, _f, _d, _d, _d]),
    "fm_batch_create": (C.c_int, [C.c_int32, C.POINTER(_h), C.POINTER(_h)]),
    "fm_batch_destroy": (None, [_h]),
    "fm_batch_last_error": (C.c_char_p, [_h]),
    "fm_batch_step": (C.c_int, [_h, C.c_int32]),
    "fm_batch_get_state": (C.c_int, [_h, C.c_int32, _f, _f, _f, _f]),
    "fm_batch_get_energies": (C.c_int, [_h, C.c_int32, C.POINTER(FmEnergies)]),
    "fm_batch_get_averages": (C.c_int, [_h, C.c_int32, C.POINTER(FmAverages)]),
    "fm_batch_set_tuning": (C.c_int, [_h, C.c_int32, C.c_int32, C.c_int32]),
    "fm_batch_tune": (C.c_int, [_h, C.c_char_p]),
    "fm_batch_set_reorder_interval": (C.c_int, [_h, C.c_int32]),
    "fm_batch_device_errors": (C.c_uint32, [_h]),
}

_LIB = None


def device_count() -> int:
    n = C.c_int32(0)
    load().fm_device_count(C.byref(n))
    return n.value


def load() -> C.CDLL:
    """Load libfemmd.so (once) and attach prototypes.  Raises if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP engine first "
            "(python -m femmd_old_amd.build)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
