"""ctypes binding of libmgp.so (the C ABI declared in include/mgp.h).

The library is built in-tree by `make -C mglet-base_amd` (or
`__graft_entry__.build()`); loading fails loudly when it is missing, there is
no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

_PKG = Path(__file__).resolve().parent.parent
LIB_PATH = _PKG / "_build" / "libmgp.so"
HEADER = _PKG.parent / "include" / "mgp.h"

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class MgpGrids(C.Structure):
    _fields_ = [("ngrid", C.c_int32), ("minlevel", C.c_int32), ("maxlevel", C.c_int32),
                ("kk", _i32p), ("jj", _i32p), ("ii", _i32p), ("level", _i32p),
                ("nbrgrid", _i32p), ("itypbc", _i32p), ("ip3", _i64p), ("ip1x", _i64p),
                ("ip1y", _i64p), ("ip1z", _i64p), ("iparent", _i32p), ("position", _i32p)]


class MgpConfig(C.Structure):
    _fields_ = [("ityp", C.c_int32), ("ninner", C.c_int32), ("nouter", C.c_int32),
                ("nouter_min", C.c_int32), ("epcorr", C.c_double), ("omega", C.c_double),
                ("rho", C.c_double), ("ibtype", C.c_int32), ("precision", C.c_int32),
                ("eps_rel", C.c_double), ("sor_all_levels", C.c_int32),
                ("check_every", C.c_int32), ("alfa", C.c_double),
                ("sync_by_caller", C.c_int32)]


class MgpStats(C.Structure):
    _fields_ = [("ipc", C.c_int32), ("nouter_done", C.c_int32), ("converged", C.c_int32),
                ("maxrhs0", C.c_double), ("maxrhs", C.c_double), ("solve_ms", C.c_double)]


_lib = None


class MgpError(RuntimeError):
    pass


def exported_symbols():
    """Function names declared in include/mgp.h."""
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(mgp_\w+)\(", txt, re.M)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise MgpError(f"{LIB_PATH} is not built; run `make -C mglet-base_amd` "
                       "or __graft_entry__.build()")
    # One HIP runtime per process: when torch is present its runtime is
    # loaded first, and libmgp's libamdhip64.so.7 dependency resolves to the
    # already-loaded copy (same soname) instead of mapping a second runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    sig = {
        "mgp_init": [C.POINTER(MgpGrids), C.POINTER(MgpConfig), _dp, _dp, _dp, _dp, _dp, _dp,
                     _dp, C.POINTER(vp)],
        "mgp_init_mask": [C.POINTER(MgpGrids), C.POINTER(MgpConfig), _dp, _dp, _dp, _dp, _dp,
                          _dp, C.POINTER(C.c_int8), C.POINTER(vp)],
        "mgp_mgpoisl": [vp, _dp, _dp, _dp, _dp, C.c_double, C.c_int32, C.c_int32,
                        C.POINTER(MgpStats)],
        "mgp_mgpoisl_device": [vp, vp, vp, vp, vp, C.c_double, C.c_int32, C.c_int32,
                               C.POINTER(MgpStats)],
        "mgp_finish": [vp],
        "mgp_set_iterations": [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                               C.c_int32],
        "mgp_set_stream": [vp, vp],
        "mgp_get_field": [vp, C.c_char_p, _dp],
        "mgp_get_field_grid": [vp, C.c_char_p, C.c_int32, _dp],
        "mgp_connect_device": [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                               C.POINTER(vp), C.c_int32],
        "mgp_divcal_device": [vp, vp, vp, vp, vp, C.c_double],
        "mgp_divcal_device_noalloc": [vp, vp, vp, vp, vp, C.c_double],
        "mgp_set_sdiv": [vp, _dp],
        "mgp_bound_layout": [vp, _i64p, _i64p],
        "mgp_set_bound": [vp, _dp, _dp, _dp, _dp],
        "mgp_time_kernels": [vp, C.c_int32, _dp],
        "mgp_plan_shape": [vp, C.c_int32, _i32p, _dp],
        "mgp_plan_shape_backward": [vp, _i32p],
        "mgp_tstle4_device": [vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double, C.c_int32, _dp],
        "mgp_rkstep_device": [vp, vp, vp, vp, C.c_double, C.c_double],
        "mgp_flow_bc_device": [vp, C.c_int32, vp, vp, vp, vp],
        "mgp_par_ftoc_device": [vp, C.c_int32, vp, vp, vp, C.c_int32],
        "mgp_info": [vp, _i64p, _i64p],
        "mgp_factor_classes": [vp, _i32p, _i64p],
        "mgp_get_history": [vp, _dp, C.c_int32, _i32p],
        "mgp_adjoint": [vp, _dp, _dp, _dp, C.c_double, C.POINTER(MgpStats)],
        "mgp_adjoint_device": [vp, vp, vp, vp, C.c_double, C.POINTER(MgpStats)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.mgp_last_error.restype = C.c_char_p
    lib.mgp_last_error.argtypes = []
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().mgp_last_error().decode()
        raise MgpError(f"{what} failed (status {rc}): {msg}")
