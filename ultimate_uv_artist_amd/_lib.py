"""ctypes binding of libuvart.so (the C ABI declared in include/uvart.h).

The library is the only compute path: when it is not built, every entry
point raises ``LibraryMissingError`` instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import threading
from pathlib import Path

LIB_PATH = Path(os.environ.get("UVART_LIB", Path(__file__).with_name("libuvart.so")))
HEADER = Path(__file__).resolve().parent.parent / "include" / "uvart.h"

UVA_OK = 0
UVA_EINVAL = -1
UVA_EHIP = -2
UVA_ENOMEM = -3
UVA_ENOBOUNDARY = -5

UVA_PLAN_BATCH = 1
UVA_PLAN_NO_WAVE = 2
UVA_PLAN_NO_TEAMS = 4
UVA_PLAN_KEEP_ORDER = 8
UVA_PLAN_NO_COARSE = 16
UVA_MODE_REFINE = 0
UVA_MODE_BATCH = 1
UVA_MODE_HIER = 2
UVA_PREC_MIXED = 0
UVA_PREC_F32 = 1
UVA_PREC_F64 = 2


class LibraryMissingError(RuntimeError):
    """libuvart.so is not built (run ``python -m ultimate_uv_artist_amd.build``)."""


class UvaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"uvart error {code}: {msg}")
        self.code = code


class PlanInfo(C.Structure):
    _fields_ = [("n_islands", C.c_int64), ("n_vertices", C.c_int64), ("n_triangles", C.c_int64),
                ("nnz", C.c_int64), ("nnz_true", C.c_int64), ("device_bytes", C.c_int64),
                ("n_group_tasks", C.c_int32), ("n_wave_tasks", C.c_int32), ("n_teams", C.c_int32),
                ("reserved", C.c_int32)]


class IslandStats(C.Structure):
    _fields_ = [("cg_iters", C.c_int32), ("cg_iters_max", C.c_int32), ("arap_iters", C.c_int32),
                ("converged", C.c_int32), ("flips0", C.c_int32), ("use_ls", C.c_int32), ("status", C.c_int32),
                ("members", C.c_int32), ("rel_resid", C.c_double), ("energy", C.c_double), ("step", C.c_double),
                ("busy_ms", C.c_double)]


class ArapOpts(C.Structure):
    _fields_ = [("iters", C.c_int32), ("line_search", C.c_int32), ("mode", C.c_int32), ("precision", C.c_int32),
                ("tol", C.c_double), ("cg_tol", C.c_double), ("cg_max_iters", C.c_int32), ("use_graph", C.c_int32),
                ("iters_folded", C.c_int32)]


class SlimOpts(C.Structure):
    _fields_ = [("iters", C.c_int32), ("cg_max_iters", C.c_int32), ("tol", C.c_double), ("cg_tol", C.c_double)]


_vp = C.c_void_p
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)

SIGNATURES = {
    "uva_version": (C.c_int, []),
    "uva_last_error": (C.c_char_p, []),
    "uva_struct_size": (C.c_int, [C.c_int32]),
    "uva_plan_create": (C.c_int, [C.POINTER(_vp), C.c_int64, _i64p, _i64p, _dp, _i64p, _i64p, _dp, _dp, _dp,
                                  C.c_uint32, _vp]),
    "uva_plan_create_matrix": (C.c_int, [C.POINTER(_vp), C.c_int64, _i64p, _i64p, _dp, _vp]),
    "uva_plan_destroy": (C.c_int, [_vp]),
    "uva_plan_get_info": (C.c_int, [_vp, C.POINTER(PlanInfo)]),
    "uva_plan_get_stats": (C.c_int, [_vp, C.POINTER(IslandStats), _vp]),
    "uva_plan_set_scale": (C.c_int, [_vp, _vp, _vp]),
    "uva_plan_set_pins": (C.c_int, [_vp, C.POINTER(C.c_int64), C.c_int64, C.c_double, _vp]),
    "uva_step_apply": (C.c_int, [_vp, _vp, _vp, C.c_int32, C.c_int32, _vp]),
    "uva_plan_pcg": (C.c_int, [_vp, _vp, _vp, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.POINTER(C.c_int32),
                               _vp]),
    "uva_overlap_flags": (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, _vp, C.c_double, C.c_int32, C.c_int32, _vp, _vp]),
    "uva_overlap_flags_noalloc": (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, _vp, C.c_double, C.c_int32, C.c_int32,
                                            _vp, _vp]),
    "uva_raster_last": (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, C.c_int32, _vp, _vp]),
    "uva_jacobi_pcg": (C.c_int, [C.c_int64, _i64p, _i64p, _dp, _vp, _vp, C.c_int32, C.c_int32, C.c_double,
                                 C.POINTER(C.c_int32), _vp]),
    "uva_csr_create": (C.c_int, [C.POINTER(_vp), C.c_int64, _i64p, _i64p, _dp, _vp]),
    "uva_csr_destroy": (C.c_int, [_vp]),
    "uva_csr_pcg": (C.c_int, [_vp, _vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                              C.POINTER(C.c_int32), _vp]),
    "uva_plan_tune": (C.c_int, [_vp, _vp]),
    "uva_plan_get_tuning": (C.c_int, [_vp, _dp]),
    "uva_plan_set_tuning": (C.c_int, [_vp, _dp, _vp]),
    "uva_lscm": (C.c_int, [_vp, _vp, _i64p, C.c_double, C.c_int32, _vp]),
    "uva_arap": (C.c_int, [_vp, _vp, C.POINTER(ArapOpts), _vp]),
    "uva_arap_profile": (C.c_int, [_vp, _vp, C.POINTER(ArapOpts), _dp, C.c_int32, _vp]),
    "uva_slim": (C.c_int, [_vp, _vp, C.POINTER(SlimOpts), _vp]),
    "uva_energy_fwd": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp, _vp]),
    "uva_energy_bwd": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp, _vp, _vp, _vp]),
    "uva_adjoint_cg": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int32, C.c_int32, _vp]),
    "uva_arap_step_bwd": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "uva_local_rhs": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp]),
    "uva_cg_solve": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int32, C.c_int32, _vp]),
    "uva_max_step": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int32, _vp]),
    "uva_energy": (C.c_int, [_vp, _vp, C.c_int32, _vp, _vp]),
    "uva_signed_areas": (C.c_int, [_vp, _vp, _vp, _vp]),
    "uva_flips_per_chart": (C.c_int, [_vp, _vp, _vp, _vp]),
    "uva_repair_jacobi": (C.c_int, [_vp, _vp, C.c_int32, _vp]),
    "uva_orient_positive": (C.c_int, [_vp, _vp, _vp]),
    "uva_repair_seq": (C.c_int, [_dp, C.c_int64, _i64p, C.c_int64, C.c_int32]),
    "uva_chart_regroup": (C.c_int, [C.c_int64, _i64p, _i64p, C.c_int64, _i64p, C.c_int64, _i64p, C.c_int64, _dp,
                                    C.c_int64, _i64p, _i64p, _i64p, _i64p, _dp, _i64p, _i64p, _i64p]),
}

_lock = threading.Lock()
_lib = None


def header_symbols():
    """Entry points declared in include/uvart.h."""
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(uva_\w+)\s*\(", HEADER.read_text(), re.M)))


def lib():
    """Load libuvart.so once (thread-safe)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise LibraryMissingError(f"{LIB_PATH} is not built: run `python -m ultimate_uv_artist_amd.build` "
                                          "(or __graft_entry__.build())")
            L = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc, what=""):
    """Turn a status code into the exception the reference would raise."""
    if rc == UVA_OK:
        return
    msg = lib().uva_last_error().decode(errors="replace")
    if rc == UVA_ENOBOUNDARY:
        raise ValueError(msg)
    if rc == UVA_EINVAL:
        raise ValueError(f"{what}: {msg}" if what else msg)
    if rc == UVA_ENOMEM:
        raise MemoryError(msg)
    raise UvaError(rc, f"{what}: {msg}" if what else msg)


def ptr(a):
    """Pointer of a torch tensor or numpy array (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def i64p(a):
    return None if a is None else a.ctypes.data_as(_i64p)


def f64p(a):
    return None if a is None else a.ctypes.data_as(_dp)
