"""ctypes bindings of liblbmibm.so (the C ABI in include/lbmibm.h).

The library is built in-tree by ``preprocess_src_amd/build.py``; importing
this module never falls back to anything else: a missing or unloadable
library raises ``LibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblbmibm.so")

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)

F64, F32, F16 = 0, 1, 2
# adjoint buffers (LBMIBM_ADJ_*)
ADJ_POPULATIONS, ADJ_VELOCITY, ADJ_LATTICE_FORCE, ADJ_POSITIONS, ADJ_NODAL_FORCE = 0, 1, 2, 3, 4
ADJ_GRAD_BODY_FORCE, ADJ_GRAD_FORCE_SCALE = 5, 6
STATUS = {
    0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_BICG_BREAKDOWN", 4: "ERR_BICG_NOCONV",
    5: "ERR_VOLUME", 6: "ERR_TOPOLOGY", 7: "ERR_STATE", 8: "ERR_VOF",
}


class LibraryMissing(RuntimeError):
    pass


class LbmibmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("lx", C.c_int), ("ly", C.c_int), ("lz", C.c_int), ("periodic_y", C.c_int),
        ("dx", C.c_double), ("dt", C.c_double), ("cr", C.c_double), ("df", C.c_double),
        ("lambda1", C.c_double), ("p", C.c_double), ("ex", C.c_double), ("ey", C.c_double),
        ("ez", C.c_double), ("force_modulated", C.c_int), ("flowstart", C.c_int), ("freq", C.c_double),
        ("b", C.c_double), ("m", C.c_int), ("c", C.c_double), ("rgs", C.c_double), ("br", C.c_double),
        ("itrmax", C.c_int), ("residual", C.c_double), ("runge_kutta", C.c_int), ("neohookean", C.c_int),
        ("volume_correction", C.c_int), ("precision", C.c_int), ("store_macros", C.c_int),
        ("fileoutput", C.c_int), ("vof_period", C.c_int), ("sor_itrmax", C.c_int), ("sor_residual", C.c_double),
        ("tube_radius", C.c_double), ("tube_staircase", C.c_int), ("wall_motion", C.c_int),
        ("wall_velocity", C.c_double), ("bicg_warm_start", C.c_int), ("residual_floor", C.c_double), ("fp32_collide", C.c_int),
    ]


class AdhesionParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("receptor_density", "rc", "k_ad", "l0", "k_on0", "k_off0", "chi_beta",
                                          "kbt", "shear_rate", "gs")] + [
        ("limit", C.c_int), ("max_neighwall", C.c_int), ("springs_per_knob", C.c_int), ("knob", C.c_double),
        ("seed", C.c_ulonglong), ("wlc", C.c_int), ("lcont", C.c_double), ("f_wlc", C.c_double), ("ra", C.c_double),
        ("alpha1", C.c_double), ("alpha2", C.c_double),
    ]


class Cells(C.Structure):
    _fields_ = [
        ("nc", C.c_int), ("voff", c_int_p), ("eoff", c_int_p),
        ("x", c_double_p), ("y", c_double_p), ("z", c_double_p), ("ele", c_int_p),
    ]


# name -> (restype, argtypes); every declaration of include/lbmibm.h
SIGNATURES = {
    "lbmibm_autotune_report": (C.c_char_p, [C.c_void_p]),
    "lbmibm_default_params": (None, [C.POINTER(Params), C.c_int, C.c_int, C.c_int]),
    "lbmibm_create": (C.c_int, [C.POINTER(Params), C.POINTER(Cells), C.POINTER(C.c_void_p)]),
    "lbmibm_destroy": (None, [C.c_void_p]),
    "lbmibm_last_error": (C.c_char_p, [C.c_void_p]),
    "lbmibm_sizes": (None, [C.c_void_p, C.POINTER(C.c_longlong), c_int_p, c_int_p, c_int_p]),
    "lbmibm_lbm_d3q19": (C.c_int, [C.c_void_p, C.c_int]),
    "lbmibm_ibm_st1": (C.c_int, [C.c_void_p]),
    "lbmibm_ibm_st2": (C.c_int, [C.c_void_p]),
    "lbmibm_stress": (C.c_int, [C.c_void_p, C.c_int]),
    "lbmibm_volume_init": (C.c_int, [C.c_void_p]),
    "lbmibm_volume_correction": (C.c_int, [C.c_void_p, c_int_p]),
    "lbmibm_step": (C.c_int, [C.c_void_p, C.c_int]),
    "lbmibm_time_kernels": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "lbmibm_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "lbmibm_synchronize": (C.c_int, [C.c_void_p]),
    "lbmibm_time_stages": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p, c_int_p]),
    "lbmibm_set_equilibrium": (C.c_int, [C.c_void_p] + [c_double_p] * 4),
    "lbmibm_upload_populations": (C.c_int, [C.c_void_p, c_double_p]),
    "lbmibm_download_populations": (C.c_int, [C.c_void_p, c_double_p]),
    "lbmibm_download_macros": (C.c_int, [C.c_void_p] + [c_double_p] * 8),
    "lbmibm_upload_macros": (C.c_int, [C.c_void_p] + [c_double_p] * 3),
    "lbmibm_download_lattice_force": (C.c_int, [C.c_void_p] + [c_double_p] * 3),
    "lbmibm_upload_lattice_force": (C.c_int, [C.c_void_p] + [c_double_p] * 3),
    "lbmibm_download_cells": (C.c_int, [C.c_void_p] + [c_double_p] * 10),
    "lbmibm_upload_cells": (C.c_int, [C.c_void_p] + [c_double_p] * 7),
    "lbmibm_download_membrane": (C.c_int, [C.c_void_p] + [c_double_p] * 5 + [c_int_p]),
    "lbmibm_download_volume": (C.c_int, [C.c_void_p] + [c_double_p] * 4),
    "lbmibm_set_cell_force_scale": (C.c_int, [C.c_void_p, c_double_p]),
    "lbmibm_bc_mask": (C.c_int, [C.c_void_p, c_int_p]),
    "lbmibm_write_restart": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int]),
    "lbmibm_read_restart": (C.c_int, [C.c_void_p, C.c_char_p, c_int_p, c_int_p]),
    "lbmibm_write_prof": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "lbmibm_write_cell": (C.c_int, [C.c_void_p, C.c_char_p]),
    "lbmibm_output_data": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, c_int_p]),
    "lbmibm_front_tracking": (C.c_int, [C.c_void_p, c_int_p, c_int_p]),
    "lbmibm_thinc_wlic": (C.c_int, [C.c_void_p]),
    "lbmibm_download_vf": (C.c_int, [C.c_void_p, c_double_p]),
    "lbmibm_upload_vf": (C.c_int, [C.c_void_p, c_double_p]),
    "lbmibm_adhesion_setup": (C.c_int, [C.c_void_p, C.POINTER(AdhesionParams), C.POINTER(C.c_ubyte)]),
    "lbmibm_adhesion": (C.c_int, [C.c_void_p]),
    "lbmibm_write_adhesion": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, c_int_p]),
    "lbmibm_adhesion_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_ubyte), c_int_p, c_double_p, c_int_p, c_double_p,
                                        c_int_p, c_double_p]),
    "lbmibm_set_body_force_field": (C.c_int, [C.c_void_p] + [c_double_p] * 3),
    "lbmibm_adjoint_reset": (C.c_int, [C.c_void_p]),
    "lbmibm_adjoint_get": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "lbmibm_adjoint_set": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "lbmibm_adj_lbm_d3q19": (C.c_int, [C.c_void_p, C.c_int]),
    "lbmibm_adj_ibm_st1": (C.c_int, [C.c_void_p]),
    "lbmibm_adj_ibm_st2": (C.c_int, [C.c_void_p]),
    "lbmibm_adj_stress": (C.c_int, [C.c_void_p, C.c_int]),
    "lbmibm_wall_shear": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_double_p]),
    "lbmibm_wall_shear_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_double_p]),
    "lbmibm_adjoint_wall_shear": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int] + [c_double_p] * 5),
    "lbmibm_mesh_icosphere": (C.c_int, [C.c_int, c_int_p, c_int_p, c_double_p, c_int_p]),
    "lbmibm_mesh_biconcave": (C.c_int, [C.c_int, c_double_p]),
    "lbmibm_place_cells": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p]),
    "lbmibm_build_topology": (C.c_int, [C.c_int, c_int_p, c_int_p, c_double_p, c_int_p, c_int_p, c_int_p,
                                        c_int_p, c_int_p, c_int_p, c_double_p, c_int_p, c_double_p, c_double_p]),
}

_lib = None


def lib():
    """Load liblbmibm.so; raises LibraryMissing if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} not built; run preprocess_src_amd/build.py")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:
            raise LibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "float64 C-contiguous array required"
    return a.ctypes.data_as(c_double_p)


def ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous, "int32 C-contiguous array required"
    return a.ctypes.data_as(c_int_p)


def check(code, ctx=None):
    if code != 0:
        msg = lib().lbmibm_last_error(ctx).decode() if ctx else ""
        raise LbmibmError(code, msg)
