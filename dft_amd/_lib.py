"""ctypes binding of libdft_hip.so (declarations: include/dft_hip.h).

The library is the only compute path: if it is missing or fails to load,
every entry point raises. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DFT_HIP_LIB", os.path.join(_HERE, "libdft_hip.so"))

DFT_OK, DFT_EINVAL, DFT_EUNSUPPORTED, DFT_EHIP = 0, -1, -2, -3
DFT_PLAN_UHF = 1
DFT_PLAN_NOSPLIT = 2
DFT_PLAN_SERIAL = 4
DFT_PLAN_STATIC_WORKSPACE = 8
DFT_PLAN_MFMA = 16
DFT_STATIC_SCRATCH_BYTES = 32 << 20  # include/dft_hip.h
STATS_COLS = 8  # DFT_STATS_COLS: values per class combination in dft_plan_stats
DFT_FP64, DFT_MIXED = 0, 1
MODES = {"fp64": DFT_FP64, "mixed": DFT_MIXED}

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "dft_version": (C.c_int, []),
    "dft_last_error": (C.c_char_p, []),
    "dft_plan_create": (C.c_int, [C.c_int, _i32p, _i32p, _i32p, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                  C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(_vp)]),
    "dft_plan_create_batch": (C.c_int, [C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _dp, _dp, _dp, _dp,
                                        _dp, C.c_double, C.c_double, C.c_int, C.POINTER(_vp)]),
    "dft_plan_destroy": (C.c_int, [_vp]),
    "dft_jk_graph_create": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "dft_graph_launch": (C.c_int, [_vp, _vp]),
    "dft_graph_destroy": (C.c_int, [_vp]),
    "dft_plan_tune": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, _vp]),
    "dft_plan_set_shapes": (C.c_int, [_vp, C.POINTER(C.c_int), C.c_int]),
    "dft_plan_stats": (C.c_int, [_vp, _i64p, C.c_int, C.POINTER(C.c_int), _i64p, _i64p]),
    "dft_schwarz": (C.c_int, [_vp, _dp]),
    "dft_jk": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "dft_jk_static": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "dft_eri_quartets": (C.c_int, [_vp, _i32p, C.c_int64, _dp]),
    "dft_oneelec": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "dft_ao_grid": (C.c_int, [_vp, _vp, C.c_int64, _vp, C.c_int64, _vp]),
    "dft_boys": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp]),
    "dft_syevj": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp]),
    "dft_syevj_tol": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, C.c_double, _vp]),
    "dft_syevj_ws_bytes": (C.c_int64, [C.c_int]),
    "dft_scf_gemm": (C.c_int, [C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp,
                               C.c_int, _vp, _vp, _vp]),
    "dft_scf_fock_residual": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dft_scf_orbital_matrix": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dft_syevj_ws": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, C.c_double, _vp, C.c_int64, _vp]),
    "dft_scf_diis_stage": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dft_scf_damp_control": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "dft_nuclear_prim": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int,
                                   _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dft_eri_prim": (C.c_int, [C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp,
                               _vp, _vp, _vp, _vp, C.c_int, _vp, _vp]),
    "dft_grad": (C.c_int, [_vp, _i32p, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dft_plan_profile": (C.c_int, [_vp, C.c_int]),
    "dft_plan_profile_read": (C.c_int, [_vp, _dp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
}

_lib = None


class DFTError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load the library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DFTError(f"{LIB_PATH} is not built: run `python -m dft_amd.build` "
                       "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != DFT_OK:
        msg = load().dft_last_error().decode(errors="replace")
        raise DFTError(f"libdft_hip error {rc}: {msg}")


def ptr(arr, ctype):
    """ctypes pointer into a contiguous NumPy array of the matching dtype."""
    return arr.ctypes.data_as(C.POINTER(ctype))
