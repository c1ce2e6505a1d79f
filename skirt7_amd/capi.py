"""ctypes binding of the C-ABI declared in include/skirt7_amd.h.

The shared library is built in-tree (skirt7_amd/lib/libskirt7_amd.so). Loading fails loudly when
it is missing: there is no CPU fallback for any entry point.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SKIRT7_AMD_LIB", os.path.join(_HERE, "lib", "libskirt7_amd.so"))

SG_FP64 = 0
SG_MIXED = 1
SG_TOPDOWN = 0
SG_NEIGHBOR = 1
SG_BOOKKEEPING = 2
SG_STAR_NPARAM = 16
SG_INSTR_NPARAM = 10
SG_MC_NPARAM = 8
SG_NCOUNTERS = 15

COUNTER_NAMES = ("packets", "own_crossings", "peel_crossings", "scatterings", "peeloffs", "stuck", "npp",
                 "walks", "rewalk_crossings", "fp64_recomputes", "kernel_us", "blocks", "transport_us",
                 "transport_launches", "walk_kernel_us")

_lib = None


class SkirtError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status (the reference's FatalError)."""


def _p(t):
    return ctypes.POINTER(t)


_D = _p(ctypes.c_double)
_I32 = _p(ctypes.c_int32)
_I64 = _p(ctypes.c_int64)
_U8 = _p(ctypes.c_uint8)
_U64 = _p(ctypes.c_uint64)
_VP = ctypes.c_void_p
_VPP = _p(ctypes.c_void_p)

_SIGS = {
    "sg_last_error": (ctypes.c_char_p, []),
    "sg_init": (ctypes.c_int, [ctypes.c_int]),
    "sg_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "sg_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "sg_set_pipeline": (ctypes.c_int, [ctypes.c_int]),
    "sg_set_batch_alone": (ctypes.c_int, [ctypes.c_uint64]),
    "sg_set_mixed_topdown": (ctypes.c_int, [ctypes.c_int]),
    "sg_set_walk_order": (ctypes.c_int, [ctypes.c_int]),
    "sg_set_slot_groups": (ctypes.c_int, [ctypes.c_int]),
    "sg_set_emit_batch": (ctypes.c_int, [ctypes.c_uint64]),
    "sg_set_slots": (ctypes.c_int, [ctypes.c_int64]),
    "sg_autotune": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                   _I64, _D]),
    "sg_fits_write": (ctypes.c_int, [ctypes.c_char_p, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.c_char_p]),
    "sg_grid_create_cartesian": (ctypes.c_int, [_D, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _VPP]),
    "sg_grid_create_tree": (ctypes.c_int, [ctypes.c_int, _D, _U8, ctypes.c_int64, ctypes.c_int, _VPP]),
    "sg_grid_create_voronoi": (ctypes.c_int, [_D, _D, ctypes.c_int64, _I64, _I32, _VPP]),
    "sg_grid_destroy": (ctypes.c_int, [_VP]),
    "sg_grid_num_cells": (ctypes.c_int, [_VP, _I64]),
    "sg_grid_whichcell": (ctypes.c_int, [_VP, ctypes.c_int64, _D, _I32]),
    "sg_grid_paths": (ctypes.c_int, [_VP, ctypes.c_int64, _D, ctypes.c_int, ctypes.c_int, _I32, _I32, _D]),
    "sg_dust_create": (ctypes.c_int, [_VP, _D, ctypes.c_int64, ctypes.c_int, _D, _D, _D, ctypes.c_int, _VPP]),
    "sg_dust_destroy": (ctypes.c_int, [_VP]),
    "sg_dust_create_multi": (ctypes.c_int, [_VP, ctypes.c_int, _D, ctypes.c_int64, ctypes.c_int, _D, _D, _D, _D,
                                            ctypes.c_int, _VPP]),
    "sg_dust_optical_depths": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int64, _D, ctypes.c_int, _D]),
    "sg_dust_path_lengths": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int64, _D, _D, ctypes.c_int, _D]),
    "sg_dust_read_absorption": (ctypes.c_int, [_VP, _D]),
    "sg_dust_clear_absorption": (ctypes.c_int, [_VP]),
    "sg_stars_create": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, _D, _D, ctypes.c_int, ctypes.c_double,
                                       _VPP]),
    "sg_stars_destroy": (ctypes.c_int, [_VP]),
    "sg_stars_update": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _D, _D, ctypes.c_int, ctypes.c_double]),
    "sg_sersic_table": (ctypes.c_int, [ctypes.c_double, _D, _D]),
    "sg_instruments_create": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _VPP]),
    "sg_instruments_destroy": (ctypes.c_int, [_VP]),
    "sg_instruments_update": (ctypes.c_int, [_VP, _D]),
    "sg_instruments_read": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "sg_instruments_clear": (ctypes.c_int, [_VP]),
    "sg_sim_create": (ctypes.c_int, [_VP, _VP, _VP, _D, _VPP]),
    "sg_sim_destroy": (ctypes.c_int, [_VP]),
    "sg_sim_run": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_int, _U64]),
    "sg_sim_run_static": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_int, _U64]),
    "sg_sim_run_batch": (ctypes.c_int, [ctypes.c_int, _VPP, _I32, _U64, ctypes.c_int, _U64]),
    "sg_gaussian_kernel": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, _D]),
    "sg_dust_emission_spectra": (ctypes.c_int, [_VP, _D, _D, _D, ctypes.c_int, _D, ctypes.c_double, _D, _D, _D,
                                                ctypes.c_int, _D, _D, ctypes.c_double, _D, _D]),
    "sg_sim_run_dust": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _D,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                       _U64]),
    "sg_fit_frames": (ctypes.c_int, [ctypes.c_int, _VPP, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                     _D, ctypes.c_double, ctypes.c_double, _D, _D, _D]),
    "sg_geometry_density": (ctypes.c_int, [_D, ctypes.c_int64, _D, _D]),
    "sg_tree_subdivide": (ctypes.c_int, [ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint32, _U8,
                                         ctypes.c_int64, _I64]),
    "sg_grid_sample_density": (ctypes.c_int, [_VP, _D, ctypes.c_double, ctypes.c_int, ctypes.c_uint32, _D]),
    "sg_tree_subdivide_multi": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_double,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, ctypes.c_uint32, _U8, ctypes.c_int64,
                                               _I64]),
    "sg_tree_subdivide_field_od": (ctypes.c_int, [ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _U8,
                                                  ctypes.c_int64, _I64]),
    "sg_grid_resample_density": (ctypes.c_int, [_VP, _D, ctypes.c_double, ctypes.c_int, ctypes.c_uint32]),
    "sg_ga_create": (ctypes.c_int, [ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint32, _VPP]),
    "sg_ga_destroy": (ctypes.c_int, [_VP]),
    "sg_ga_initialize": (ctypes.c_int, [_VP]),
    "sg_ga_step": (ctypes.c_int, [_VP]),
    "sg_ga_pending": (ctypes.c_int, [_VP, _I32, _D]),
    "sg_ga_set_scores": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "sg_ga_state": (ctypes.c_int, [_VP, _I32, _I32]),
    "sg_ga_best": (ctypes.c_int, [_VP, _D, _D]),
    "sg_ga_population": (ctypes.c_int, [_VP, _D, _D]),
    "sg_tree_subdivide_field": (ctypes.c_int, [ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_double, _U8, ctypes.c_int64,
                                               _I64]),
    "sg_grid_field_density": (ctypes.c_int, [_VP, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]),
    "sg_voronoi_neighbors": (ctypes.c_int, [_D, _D, ctypes.c_int64, _I64, _I32, ctypes.c_int64, _I64]),
    "sg_sph_create": (ctypes.c_int, [_D, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _VPP]),
    "sg_sph_destroy": (ctypes.c_int, [_VP]),
    "sg_sph_info": (ctypes.c_int, [_VP, _I64, _D, _D, _I64]),
    "sg_sph_cell_list": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I32, ctypes.c_int64, _I64]),
    "sg_sph_mass_in_box": (ctypes.c_int, [_VP, _D, ctypes.c_int64, _D]),
    "sg_sph_density": (ctypes.c_int, [_VP, _D, ctypes.c_int64, _D]),
    "sg_tree_subdivide_sph": (ctypes.c_int, [ctypes.c_int, _D, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_uint32, _U8, _D,
                                             ctypes.c_int64, _I64]),
    "sg_tree_subdivide_device": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int, ctypes.c_uint32, _U8, ctypes.c_int64, _I64]),
    "sg_grid_sph_density": (ctypes.c_int, [_VP, _VP, _D, ctypes.c_int64, _D]),
    "sg_knn_distance": (ctypes.c_int, [_D, ctypes.c_int64, ctypes.c_int, _D]),
    "sg_grid_sample_density_device": (ctypes.c_int, [_VP, _D, ctypes.c_double, ctypes.c_int, ctypes.c_uint32, _D]),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Loads the HIP library (once); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SkirtError("libskirt7_amd.so is not built; run __graft_entry__.build() (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().sg_last_error().decode(errors="replace")
        raise SkirtError("skirt7_amd: %s (status %d)" % (msg, status))


def dptr(a):
    return a.ctypes.data_as(_D)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def as_u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)
