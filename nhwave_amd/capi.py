"""ctypes binding of libnhwave_amd (declarations in include/nhwave_amd.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the in-tree shared object is missing or was built without a device
target.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libnhwave_amd.so"

NHW_OK = 0
NHW_EINVAL = -1
NHW_EHIP = -2
NHW_EDT = -3
NHW_ESOLVER = -4
NHW_ENOMEM = -5
NHW_ENOTFINITE = -6
NHW_FP64 = 0
NHW_MIXED = 1
NHW_SOLVER_MG = 0
NHW_SOLVER_LINE = 1

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class NhwConfig(C.Structure):
    _fields_ = [
        ("mglob", C.c_int32), ("nglob", C.c_int32), ("kglob", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double),
        ("cfl", C.c_double), ("min_dep", C.c_double),
        ("dt_ini", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("froude_cap", C.c_double), ("froude_cap_on", C.c_int32),
        ("visc", C.c_double), ("fcor", C.c_double),
        ("wind_u", C.c_double), ("wind_v", C.c_double),
        ("time_order", C.c_int32), ("non_hydro", C.c_int32), ("hllc", C.c_int32),
        ("bc_x0", C.c_int32), ("bc_xn", C.c_int32), ("bc_y0", C.c_int32),
        ("bc_yn", C.c_int32), ("bc_z0", C.c_int32),
        ("itmax", C.c_int32), ("tol", C.c_double), ("precision", C.c_int32),
        ("solver", C.c_int32),
        ("wavemaker", C.c_int32), ("amp_wave", C.c_double), ("per_wave", C.c_double),
        ("dep_wave", C.c_double), ("theta_wave", C.c_double), ("cur_wave", C.c_double),
        ("tramp", C.c_double), ("sponge_on", C.c_int32), ("sponge_w", C.c_double),
        ("sponge_e", C.c_double), ("sponge_s", C.c_double), ("sponge_n", C.c_double),
        ("viscous", C.c_int32), ("rng", C.c_int32), ("cvs", C.c_double), ("chs", C.c_double),
        ("schmidt", C.c_double), ("visc_number", C.c_double),
        ("ibot", C.c_int32), ("cd0", C.c_double), ("zob", C.c_double),
        ("turb", C.c_int32),
    ]


# WAVEMAKER names of the reference's input.txt -> NHW_WAVE_*
WAVEMAKERS = {"NONE": 0, "LEF_SOL": 1, "LEF_LIN": 2, "LEF_STK": 3, "LEF_CON": 4, "INT_IRR": 5, "FLUX_LR": 6}


FIELD_NAMES_D1 = ["Eta", "Eta0", "D", "D0", "Hc", "DelxH", "DelyH", "DelxEta", "DelyEta"]
FIELD_NAMES_I = ["Mask", "Mask_Struct", "Mask9"]
FIELD_NAMES_D2 = [
    "Hfx", "Hfx0", "EtaxL", "EtaxR", "Hfy", "Hfy0", "EtayL", "EtayR",
    "U", "V", "W", "DU", "DV", "DW", "DU0", "DV0", "DW0", "P", "Omega",
    "SourceX", "SourceY", "Wsx", "Wsy",
    "Ex", "Ey", "Fx", "Fy", "Gx", "Gy", "Hx", "Hy", "Fz", "Gz", "Hz",
    "CmuVt", "Tke", "Eps", "DTke", "DEps",
]


class NhwFields(C.Structure):
    _fields_ = ([(n, _DP) for n in FIELD_NAMES_D1] + [(n, _IP) for n in FIELD_NAMES_I]
                + [(n, _DP) for n in FIELD_NAMES_D2])


class NhwStatValues(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("mass_volume", "energy", "max_eta", "min_eta", "max_u", "max_v",
                                           "max_w")]


class NhwStats(C.Structure):
    _fields_ = [
        ("last_iters", C.c_int32), ("last_relres", C.c_double),
        ("total_iters", C.c_int64), ("solves", C.c_int64),
        ("max_iters", C.c_int32), ("not_converged", C.c_int32),
    ]


EXPORTS = {
    "nhw_last_error": (C.c_char_p, []),
    "nhw_create": (C.c_int, [C.POINTER(NhwConfig), C.POINTER(C.c_void_p)]),
    "nhw_destroy": (C.c_int, [C.c_void_p]),
    "nhw_footprint": (C.c_int, [C.POINTER(NhwConfig), C.POINTER(C.c_int64)]),
    "nhw_set_grid": (C.c_int, [C.c_void_p, _DP, _DP, _DP]),
    "nhw_upload": (C.c_int, [C.c_void_p, C.POINTER(NhwFields)]),
    "nhw_download": (C.c_int, [C.c_void_p, C.POINTER(NhwFields)]),
    "nhw_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int32]),
    "nhw_get_time": (C.c_int, [C.c_void_p, _DP, _DP, _IP]),
    "nhw_initial": (C.c_int, [C.c_void_p]),
    "nhw_estimate_dt": (C.c_int, [C.c_void_p]),
    "nhw_update_mask": (C.c_int, [C.c_void_p]),
    "nhw_update_wind": (C.c_int, [C.c_void_p]),
    "nhw_update_vars": (C.c_int, [C.c_void_p]),
    "nhw_source_terms": (C.c_int, [C.c_void_p]),
    "nhw_fluxes": (C.c_int, [C.c_void_p]),
    "nhw_eval_duvw": (C.c_int, [C.c_void_p, C.c_int32]),
    "nhw_poisson_solver": (C.c_int, [C.c_void_p]),
    "nhw_generate_coef_rhs": (C.c_int, [C.c_void_p, _DP, _DP]),
    "nhw_poisson_apply": (C.c_int, [C.c_void_p, _DP, _DP]),
    "nhw_update_wave_bc": (C.c_int, [C.c_void_p]),
    "nhw_set_internal_wavemaker": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.c_int32, C.c_int32, _DP, _DP, _DP, _DP]),
    "nhw_sponge_damping": (C.c_int, [C.c_void_p]),
    "nhw_eval_turb": (C.c_int, [C.c_void_p]),
    "nhw_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "nhw_run": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, C.POINTER(C.c_int32)]),
    "nhw_solver_stats": (C.c_int, [C.c_void_p, C.POINTER(NhwStats)]),
    "nhw_sync": (C.c_int, [C.c_void_p]),
    "nhw_kernel_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "nhw_kernel_stats": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_char_p),
                                   C.POINTER(C.c_int64), _DP, _DP]),
    "nhw_bench_copy": (C.c_int, [C.c_int64, C.c_int32, _DP]),
    "nhw_batch_create": (C.c_int, [C.POINTER(NhwConfig), C.c_int32, C.POINTER(C.c_void_p)]),
    "nhw_batch_member": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "nhw_batch_add": (C.c_int, [C.c_void_p, C.POINTER(NhwConfig), C.POINTER(C.c_void_p)]),
    "nhw_batch_remove": (C.c_int, [C.c_void_p, C.c_int32]),
    "nhw_batch_size": (C.c_int, [C.c_void_p, _IP]),
    "nhw_batch_prepare": (C.c_int, [C.c_void_p, C.c_int32]),
    "nhw_batch_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "nhw_batch_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_char_p), _IP, _DP]),
    "nhw_batch_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "nhw_statistics": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nhw_snapshot_begin": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nhw_snapshot_wait": (C.c_int, [C.c_void_p]),
    "nhw_snapshot_static": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nhw_checkpoint_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "nhw_checkpoint_load": (C.c_int, [C.c_void_p, C.c_char_p]),
    "nhw_batch_destroy": (C.c_int, [C.c_void_p]),
}

_lib = None


class NhwError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed ({code}): {last_error()}")
        self.code = code


def load() -> C.CDLL:
    """Loads the in-tree library once; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m nhwave_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    if _lib is None:
        return ""
    msg = _lib.nhw_last_error()
    return msg.decode() if msg else ""


def check(code: int, what: str) -> None:
    if code != NHW_OK:
        raise NhwError(code, what)
