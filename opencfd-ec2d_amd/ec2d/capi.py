"""ctypes binding of include/ec2d.h (the C-ABI of libec2d_hip.so).

The library is the product: there is no Python or CPU fallback.  Loading
fails loudly if the built library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libec2d_hip.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "ec2d.h")

EC2D_U, EC2D_RES, EC2D_DT, EC2D_DU = 1, 2, 3, 4
EC2D_LCI, EC2D_LCJ, EC2D_LVI, EC2D_LVJ = 5, 6, 7, 8
EC2D_AMU, EC2D_AMU_T = 9, 10
EC2D_UN, EC2D_QF, EC2D_DELTU = 23, 24, 25
EC2D_UN1 = 31
EC2D_DW = 22
EC2D_VOL, EC2D_SI, EC2D_SJ, EC2D_NI1, EC2D_NI2, EC2D_NJ1, EC2D_NJ2 = 11, 12, 13, 14, 15, 16, 17
EC2D_X1, EC2D_Y1 = 20, 21
EC2D_CLOSURE_MUT, EC2D_CLOSURE_MASK = 40, 41
CLOSURE_NPARAM = 6 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1  # W1 b1 W2 b2 W3 b3 W4 b4
CLOSURE_NNORM = 24  # mean, inv_std, lo, hi of the 6 features


class Control(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Ma", "Re", "gamma", "AoA", "Pr", "Cv", "Cp", "t_end", "dt_global", "CFL",
                                          "dtmax", "dtmin", "P_outlet", "T_inf", "Twall", "vt_inf", "Kt_inf",
                                          "Wt_inf")] + \
               [(n, C.c_int) for n in ("If_viscous", "Iflag_turbulence_model", "Iflag_init", "Iflag_local_dt",
                                       "Time_Method", "Iflag_Scheme", "Iflag_Flux", "IFlag_Reconstruction",
                                       "Num_Mesh", "Nvar")] + [("Pre_Step_Mesh", C.c_int * 3)] + \
               [("Nstep_Inner_Limit", C.c_int), ("Res_Inner_Limit", C.c_double)]


# exported entry points and their argument types
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i = C.c_int
SIGNATURES = {
    "ec2d_init": [],
    "ec2d_finalize": [],
    "ec2d_last_error": [C.c_char_p, _i],
    "ec2d_case_create": [C.POINTER(Control)],
    "ec2d_define_block": [_i, _i, _i, _i, _i, _i, _ip, _i] + [_dp] * 13,
    "ec2d_commit": [],
    "ec2d_put_array": [_i, _i, _i, _i, _dp],
    "ec2d_get_array": [_i, _i, _i, _i, _dp],
    "ec2d_init_flow": [],
    "ec2d_comput_residual_one_mesh": [_i],
    "ec2d_du_lu_sgs_2d": [_i, _i, C.c_double],
    "ec2d_du_lu_sgs_2d_noalloc": [_i, _i, C.c_double],
    "ec2d_update_u": [_i],
    "ec2d_boundary_condition_onemesh": [_i],
    "ec2d_update_buffer_onemesh": [_i],
    "ec2d_force_vt_kw": [_i],
    "ec2d_ns_time_advance": [_i],
    "ec2d_ns_multigrid": [],
    "ec2d_interpolation2h": [_i, _i, _i],
    "ec2d_prolong_u": [_i, _i, _i],
    "ec2d_comput_force_function": [_i],
    "ec2d_add_force_function": [_i],
    "ec2d_comput_delt_u": [_i],
    "ec2d_set_un": [_i],
    "ec2d_comput_new_u": [_i],
    "ec2d_run": [_i],
    "ec2d_run_to_drop": [C.c_double, _i, _i, _ip, _ip, _dp],
    "ec2d_res_norms": [_i, _i, _dp, _dp],
    "ec2d_time": [_i, _i, _ip, _dp],
    "ec2d_dual_inner": [_ip],
    "ec2d_set_time": [_i, _i, _i, C.c_double],
    "ec2d_get_run_counters": [C.POINTER(C.c_longlong)],
    "ec2d_set_run_counters": [C.POINTER(C.c_longlong)],
    "ec2d_synchronize": [],
    "ec2d_num_cells": [_i, _i],
    "ec2d_timing": [_dp, _dp, _dp, _dp],
    "ec2d_set_profile": [_i],
    "ec2d_set_wait_budget": [_i],
    "ec2d_timing_level": [_i, _dp, _dp, _dp, _ip],
    "ec2d_kernel_time": [_i, _i, _dp, _ip],
    "ec2d_force_coefficients": [_i, _dp, _dp, _dp],
    "ec2d_autotune": [C.c_char_p, _i, C.c_char_p, _i],
    "ec2d_closure_collect": [_ip],
    "ec2d_closure_train": [_i, _i, C.c_double, C.c_uint, _dp, _dp],
    "ec2d_closure_gradient": [_i, C.POINTER(C.c_float), _dp],
    "ec2d_closure_get_params": [C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "ec2d_closure_set_params": [C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "ec2d_closure_infer": [],
    "ec2d_closure_enable": [_i],
    "ec2d_lusgs_shape": [_i, _i, _i],
    "ec2d_closure_relaxation": [C.c_double],
    "ec2d_closure_monitor": [C.c_double, _ip],
    "ec2d_set_tile_rows": [_i],
    "ec2d_set_tile_shape": [_i, _i],
}

_lib = None


class Ec2dError(RuntimeError):
    pass


def load():
    """Load the built library (no GPU needed to load it)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Ec2dError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_longlong if name == "ec2d_num_cells" else C.c_int
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(512)
    load().ec2d_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str = ""):
    if rc != 0:
        raise Ec2dError(f"{what}: {last_error()} (status {rc})")
    return rc


def call(name: str, *args):
    return check(getattr(load(), name)(*args), name)
