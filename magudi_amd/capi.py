"""ctypes binding of the C ABI in include/magudi_hip.h (libmagudi_hip.so).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmagudi_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "magudi_hip.h")

MGD_OK = 0
SOLUTION_LIMIT = 1
FORWARD, ADJOINT = 1, -1
CONSERVED, ADJOINT_VARS, RHS, INITIAL_STATE, FINAL_STATE, SCRATCH = 0, 1, 2, 4, 5, 6
FIELD_TARGET_STATE, FIELD_CONTROL_MOLLIFIER, FIELD_TARGET_MOLLIFIER = 0, 1, 2
OP_FIRST, OP_ADJOINT_FIRST, OP_DISS, OP_DISS_T, OP_SECOND = 0, 1, 2, 3, 4
OPT_FORCE_PLAIN, OPT_TILE_SHAPE, OPT_RESIDENT_STATES, OPT_WINDOW_CACHE, OPT_Z_CHUNK = 1, 2, 3, 4, 5
OPT_MARCH_TILE, OPT_STEP_STORE = 6, 7
TERMINAL_KINETIC_ENERGY = 0
KERNEL_FWD_STAGE, KERNEL_ADJ_STAGE, KERNEL_COPY, KERNEL_RHS_FORWARD, KERNEL_RHS_ADJOINT = 0, 1, 2, 3, 4
KERNEL_FWD_MARCH3D = 5


class RegionDesc(C.Structure):
    _fields_ = [("nd", C.c_int), ("n", C.c_int * 3), ("periodicity", C.c_int * 3),
                ("periodic_length", C.c_double * 3), ("scheme", C.c_int), ("viscous", C.c_int),
                ("dissipation", C.c_int), ("composite_dissipation", C.c_int),
                ("gamma", C.c_double), ("reynolds", C.c_double), ("prandtl", C.c_double),
                ("power_law", C.c_double), ("bulk_ratio", C.c_double),
                ("dissipation_amount", C.c_double), ("n_instances", C.c_int)]


class PatchDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("normal_dir", C.c_int), ("extent", C.c_int * 6),
                ("p", C.c_double * 4)]


class MagudiError(RuntimeError):
    pass


_lib = None


def declared_symbols():
    """Entry points declared in include/magudi_hip.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mgd_[a-z_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MagudiError(f"{LIB_PATH} is missing: build it with `python -m magudi_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.mgd_last_error.restype = C.c_char_p
    sigs = {
        "mgd_device_count": [C.POINTER(C.c_int)],
        "mgd_region_create": [C.POINTER(RegionDesc), dp, C.POINTER(vp)],
        "mgd_region_destroy": [vp],
        "mgd_region_add_patch": [vp, C.POINTER(PatchDesc)],
        "mgd_region_set_field": [vp, C.c_int, dp],
        "mgd_region_add_source": [vp, dp, C.c_double, C.c_double, C.c_double, C.c_double],
        "mgd_region_finalize": [vp],
        "mgd_state_put": [vp, C.c_int, C.c_int, dp],
        "mgd_state_get": [vp, C.c_int, C.c_int, dp],
        "mgd_compute_rhs": [vp, C.c_int, C.c_double, dp, C.c_double],
        "mgd_update_adjoint_forcing": [vp, C.c_int],
        "mgd_functional": [vp, dp],
        "mgd_run_forward": [vp, C.c_double, C.c_double, C.c_int, dp, dp, C.c_int],
        "mgd_run_adjoint": [vp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, dp, dp, C.c_int],
        "mgd_group_create": [C.POINTER(vp), C.c_int, C.POINTER(vp)],
        "mgd_group_destroy": [vp],
        "mgd_group_run_adjoint": [vp, C.c_double, dp, C.c_int, C.c_int, C.POINTER(dp),
                                  C.POINTER(dp), C.POINTER(dp), C.POINTER(dp)],
        "mgd_stencil_apply": [vp, C.c_int, C.c_int, dp, C.c_int],
        "mgd_region_get_grid": [vp, C.c_int, dp],
        "mgd_last_timings": [vp, dp, C.POINTER(C.c_int)],
        "mgd_sync": [vp],
        "mgd_region_set_option": [vp, C.c_int, C.c_int],
        "mgd_region_get_tile_shape": [vp, C.POINTER(C.c_int)],
        "mgd_region_get_march_shape": [vp, C.POINTER(C.c_int)],
        "mgd_region_get_checkpoint_plan": [vp, C.POINTER(C.c_int)],
        "mgd_time_kernel": [vp, C.c_int, C.c_int, dp, dp],
        "mgd_kernel_timer": [vp, C.c_int, C.c_int],
        "mgd_kernel_timer_read": [vp, C.POINTER(C.c_long), C.POINTER(C.c_long), dp],
        "mgd_state_copy": [vp, C.c_int, C.c_int],
        "mgd_state_lincomb": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_double, C.c_int, C.c_int],
        "mgd_inner_product": [vp, C.c_int, C.c_int, dp],
        "mgd_inner_product_noalloc": [vp, C.c_int, C.c_int, vp],
        "mgd_region_set_time_offsets": [vp, dp],
        "mgd_run_forward_checkpointed": [vp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp],
        "mgd_run_adjoint_checkpointed": [vp, C.c_int, dp, dp],
        "mgd_terminal_cost": [vp, C.c_int, dp],
        "mgd_rk4_substep": [vp, C.c_int, C.c_int, C.c_double, C.c_double, dp],
        "mgd_checkpoint_migrate": [vp, C.c_int, C.c_int],
        "mgd_control_gradient": [vp, dp, dp],
        "mgd_adjoint_initial_condition": [vp, C.c_int],
        "mgd_region_set_solution_limits": [vp, C.c_int, dp],
        "mgd_check_solution_limits": [vp, C.POINTER(C.c_int)],
        "mgd_time_step_size": [vp, C.c_double, dp],
        "mgd_cfl": [vp, C.c_double, dp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != MGD_OK:
        msg = lib().mgd_last_error()
        raise MagudiError(f"magudi_hip error {rc}: {msg.decode() if msg else ''}")


def dptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))


def device_count() -> int:
    n = C.c_int(0)
    check(lib().mgd_device_count(C.byref(n)))
    return n.value
