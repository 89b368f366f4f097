"""ctypes binding of libhamsx.so (the C-ABI declared in include/hamsx.h).

The library is built in-tree (``make -C hams-mrel_amd``); importing this
module fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhamsx.so")

HAMSX_OK = 0
HAMSX_EARG = -1
HAMSX_EDEVICE = -2
HAMSX_ENOMEM = -3
HAMSX_ESTATE = -4
HAMSX_ENOCONV = -5
HAMSX_NK = 200
FREQ_LIMIT = 0  # HAMSX_FREQ_LIMIT: hamsx_wave_properties of the zero / infinite frequency limits
HAMSX_NOALLOC_MAX_N = 32768  # largest order hamsx_factor_noalloc takes
PREC_FP64 = 0
PREC_MIXED = 1
SOLVER_LU = 0
SOLVER_GMRES = 1
FIELD_POTENTIAL = 0
FIELD_PRESSURE = 1
FIELD_ELEVATION = 2
FIELD_AUTO = 3
OPT_LU_LOOKAHEAD = 1
OPT_LU_SELECT_ROWS = 2
OPT_LU_OUTER_MIN_N = 3
OPT_GRAPHS = 4
OPT_FP32_FARFIELD = 5
OPTIONS = {"lu_lookahead": OPT_LU_LOOKAHEAD, "lu_select_rows": OPT_LU_SELECT_ROWS,
           "lu_outer_min_n": OPT_LU_OUTER_MIN_N, "graphs": OPT_GRAPHS, "fp32_farfield": OPT_FP32_FARFIELD}

#: every symbol include/hamsx.h declares, with (restype, argtypes)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
SIGNATURES = {
    "hamsx_version": (C.c_char_p, []),
    "hamsx_last_error": (C.c_char_p, []),
    "hamsx_device_count": (C.c_int, []),
    "hamsx_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "hamsx_destroy": (C.c_int, [_vp]),
    "hamsx_set_precision": (C.c_int, [_vp, C.c_int]),
    "hamsx_set_option": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hamsx_get_option": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "hamsx_panel_geometry": (C.c_int, [C.c_int, _dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp]),
    "hamsx_wave_properties": (C.c_int, [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]),
    "hamsx_set_mesh": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip]),
    "hamsx_set_lid": (C.c_int, [_vp, C.c_int, _dp]),
    "hamsx_set_symmetry": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hamsx_set_frequency": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, C.c_int]),
    "hamsx_set_waves": (C.c_int, [_vp, C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "hamsx_assemble": (C.c_int, [_vp]),
    "hamsx_factor": (C.c_int, [_vp]),
    "hamsx_factor_noalloc": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp]),
    "hamsx_solve_radiation": (C.c_int, [_vp, _dp]),
    "hamsx_solve_diffraction": (C.c_int, [_vp, _dp]),
    "hamsx_forces": (C.c_int, [_vp, _dp, _dp, _dp]),
    "hamsx_solve_frequency": (C.c_int, [_vp, C.c_double, C.c_double, _dp, _dp, _dp]),
    "hamsx_submit_frequency": (C.c_int, [_vp, C.c_double, C.c_double]),
    "hamsx_submit_frequency_group": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(C.c_double), C.c_double]),
    "hamsx_query": (C.c_int, [_vp]),
    "hamsx_collect": (C.c_int, [_vp, _dp, _dp, _dp]),
    "hamsx_queue_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "hamsx_queue_destroy": (C.c_int, [_vp]),
    "hamsx_set_queue": (C.c_int, [_vp, _vp, C.c_int]),
    "hamsx_queue_wait": (C.c_int, [_vp, C.c_double, C.POINTER(C.c_int)]),
    "hamsx_graph_stats": (C.c_int, [_vp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "hamsx_get_matrix": (C.c_int, [_vp, C.c_int, _dp, _ip]),
    "hamsx_set_solver": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int]),
    "hamsx_set_refinement": (C.c_int, [_vp, C.c_int, C.c_double]),
    "hamsx_set_hmatrix": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]),
    "hamsx_hmatrix_stats": (C.c_int, [_vp, _ip, _ip, _dp, _dp]),
    "hamsx_hmatrix_matvec": (C.c_int, [_vp, _dp, C.c_int, _dp]),
    "hamsx_hmatrix_partition": (C.c_int, [C.c_int, _dp, C.c_int, C.c_double, _ip, C.c_int, _ip, _ip]),
    "hamsx_refine_stats": (C.c_int, [_vp, _ip, _dp]),
    "hamsx_set_field_points": (C.c_int, [_vp, C.c_int, _dp]),
    "hamsx_field_values": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip]),
    "hamsx_set_body_matrices": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "hamsx_motions": (C.c_int, [_vp, _dp, _ip]),
    "hamsx_set_profiling": (C.c_int, [_vp, C.c_int]),
    "hamsx_stage_times": (C.c_int, [_vp, C.c_int, _dp, _ip]),
}
STAGES = ("assemble", "rhs", "factor", "solve", "forces", "asm_linton", "asm_pidcock", "asm_eigene", "asm_eigen",
          "field")

_lib = None


class HamsxError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libhamsx.so not built at {LIB_PATH}; run `make -C hams-mrel_amd` "
                          "(or __graft_entry__.build()) first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, where: str, allow_info: bool = False) -> int:
    """Raise on negative codes; positive INFO values are returned (LAPACK style)."""
    if rc < 0 or (rc > 0 and not allow_info):
        msg = load().hamsx_last_error().decode(errors="replace")
        raise HamsxError(rc, where, msg)
    return rc


def ptr(a, ctype=C.c_double):
    """Pointer to the first element of a contiguous array (C or Fortran order);
    complex arrays are passed as interleaved doubles."""
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        raise ValueError("array must be contiguous")
    return C.cast(a.ctypes.data, C.POINTER(ctype))
