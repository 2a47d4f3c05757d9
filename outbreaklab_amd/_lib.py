"""ctypes binding of liboutbreak_hip.so (the C ABI in include/outbreak_hip.h).

The product path has no fallback: if the gfx950 library is missing or fails
to load, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OBL_LIB", os.path.join(_PKG, "liboutbreak_hip.so"))

_lib = None


class OblError(RuntimeError):
    pass


class SurrStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("steps", "points", "fallbacks", "overflow", "judged")] + [
        ("d2", C.c_double * 8), ("r2", C.c_double * 8)] + [
        (n, C.c_double) for n in ("cc_d2", "cc_r2", "cc_maxdiff", "cc_max", "dbz_d2", "dbz_cols", "dbz_maxdiff",
                                  "pr_d2", "pr_r2", "pr_sum_s", "pr_sum_p")]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dz", C.c_double),
                ("pad", C.c_int)]


# name -> (restype, argtypes)
_SIGS = {
    "obl_abi_version": (C.c_int, []),
    "obl_grid_bytes": (C.c_double, [C.c_void_p]),
    "obl_last_error": (C.c_char_p, [C.c_void_p]),
    "obl_global_error": (C.c_char_p, []),
    "obl_create": (C.c_int, [C.POINTER(Grid), C.POINTER(C.c_void_p)]),
    "obl_destroy": (None, [C.c_void_p]),
    "obl_set_base_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_set_boundaries": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_set_ci": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_set_surface": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "obl_set_prep_field": (C.c_int, [C.c_void_p, C.c_int]),
    "obl_set_vortex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]),
    "obl_vortex_table": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_void_p, C.c_size_t, C.c_void_p]),
    "obl_reset": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint]),
    "obl_maps_request": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "obl_maps_collect": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "obl_radar_create": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.POINTER(C.c_void_p)]),
    "obl_radar_destroy": (None, [C.c_void_p]),
    "obl_radar_set_tornadoes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_radar_cut": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_radar_cut_into": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "obl_radar_geometry": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "obl_radar_readback": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "obl_radar_collect": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "obl_add_bubble": (C.c_int, [C.c_void_p] + [C.c_double] * 6),
    "obl_upload": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_size_t]),
    "obl_download": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_size_t]),
    "obl_field_ptr": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "obl_sync": (C.c_int, [C.c_void_p]),
    "obl_set_solver_mode": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
    "obl_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "obl_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_int]),
    "obl_dump_enable": (C.c_int, [C.c_void_p, C.c_uint, C.c_float, C.c_float]),
    "obl_dump_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_void_p)]),
    "obl_mlp_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "obl_mlp_destroy": (None, [C.c_void_p]),
    "obl_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "obl_mlp_nparams": (C.c_int, []),
    "obl_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "obl_mlp_set_scales": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "obl_mlp_load_rows": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "obl_mlp_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint,
                                C.c_int, C.POINTER(C.c_float)]),
    "obl_mlp_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_void_p,
                               C.c_void_p]),
    "obl_mlp_set_dead_zone": (C.c_int, [C.c_void_p, C.c_float]),
    "obl_mlp_set_hand_min": (C.c_int, [C.c_void_p, C.c_float]),
    "obl_mlp_infer": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "obl_set_surrogate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint]),
    "obl_surrogate_stats": (C.c_int, [C.c_void_p, C.POINTER(SurrStats), C.c_int]),
    "obl_poisson_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    "obl_poisson_destroy": (None, [C.c_void_p]),
    "obl_poisson_set_density": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "obl_poisson_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p]),
    "obl_poisson_solve_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double]),
    "obl_poisson_last_error": (C.c_char_p, [C.c_void_p]),
    "obl_poisson_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "obl_poisson_profile_read": (C.c_int, [C.c_void_p, C.c_void_p]),
}

FIELD = {"V": 0, "Q1": 1, "Q2": 2, "Vt": 3, "Q1t": 4, "Q2t": 5, "aux": 6, "phT": 7, "lsfA": 8, "lsfB": 9,
         "ciMap": 10, "phi": 11, "rhs": 12, "bs": 13, "sfc": 14, "Vp": 15}
STEP_LSF = 1
STEP_JUDGE = 2
STEP_FILT = 4
GEMM = {"fp32": 0, "bf16x3": 1, "bf16x6": 2}
THOMAS = {"fp32": 0, "fp64": 1, "mixed": 2}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OblError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc for gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def symbols():
    return list(_SIGS)


def check(rc: int, handle=None, poisson=False, mlp=False):
    if rc == 0:
        return
    L = lib()
    if mlp:
        msg = L.obl_mlp_last_error(handle)
    elif handle is None:
        msg = L.obl_global_error()
    elif poisson:
        msg = L.obl_poisson_last_error(handle)
    else:
        msg = L.obl_last_error(handle)
    raise OblError((msg or b"unknown error").decode(errors="replace"))
