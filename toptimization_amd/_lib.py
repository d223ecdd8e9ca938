"""ctypes binding of libtopt.so (declarations mirror include/topt.h).

The library is the only compute path: if it is missing or fails to load,
every entry point raises; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libtopt.so"

F32, F64 = 0, 1
PC_NONE, PC_JACOBI, PC_ELEMBLOCK, PC_BLOCK2, PC_MG = 0, 1, 2, 3, 4
PRECONDITIONER_IDS = {"none": PC_NONE, "jacobi": PC_JACOBI, "block_jacobi": PC_ELEMBLOCK,
                      "block_jacobi_2x2": PC_BLOCK2, "multigrid": PC_MG}
FILTER_DENSITY, FILTER_DCAVG, FILTER_SENS, FILTER_TRANSPOSE = 0, 1, 2, 3
ST_RUN, ST_CONV, ST_BREAK, ST_MAXIT, ST_ZERO, ST_DONE = 0, 1, 2, 3, 4, 5

BUF_NAMES = ["u", "f", "fix", "rho", "rhof", "dc", "dcf", "Hs", "x", "r", "z", "p0", "p1", "Ap",
             "pre", "pre2", "s", "s64", "taps", "cg", "oc", "part", "opart", "ctr", "mg"]
BUF = {n: i for i, n in enumerate(BUF_NAMES)}
NOALLOC_SCRATCH_BYTES = 4 << 20   # TOPT_NOALLOC_SCRATCH_BYTES


class TProblem(C.Structure):
    _fields_ = [("nelx", C.c_int32), ("nely", C.c_int32), ("precond", C.c_int32),
                ("max_cg_iter", C.c_int32), ("E_min", C.c_double), ("penal", C.c_double),
                ("cg_tol", C.c_double), ("volfrac", C.c_double), ("move", C.c_double),
                ("rmin", C.c_double), ("fnorm2", C.c_double), ("delta", C.c_double),
                ("Ke", C.c_double * 64), ("ws", C.c_void_p), ("oc_tol", C.c_double),
                ("oc_max_iter", C.c_int32), ("pad", C.c_int32)]


class TGrid(C.Structure):
    _fields_ = [("nelx", C.c_int32), ("nely", C.c_int32), ("E_min", C.c_double),
                ("penal", C.c_double), ("Ke", C.c_double * 64)]


class TGrid3(C.Structure):
    _fields_ = [("nelx", C.c_int32), ("nely", C.c_int32), ("nelz", C.c_int32), ("pad", C.c_int32),
                ("E_min", C.c_double), ("penal", C.c_double), ("Ke", C.c_void_p)]


class TIterStats(C.Structure):
    _fields_ = [("compliance", C.c_double), ("volume", C.c_double), ("change", C.c_double),
                ("lam", C.c_double), ("n_cg", C.c_int32), ("cg_status", C.c_int32),
                ("oc_steps", C.c_int32), ("refines", C.c_int32), ("iter", C.c_int32), ("done", C.c_int32)]


_vp, _i32, _i64, _d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
_SIGS = {
    "topt_last_error": (C.c_char_p, []),
    "topt_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "topt_problem_layout": (C.c_int, [_i32, _i32, _i32, _i32, _d, C.POINTER(_i64), C.POINTER(_i64)]),
    "topt_problem_init": (C.c_int, [C.POINTER(TProblem), _i32, _vp]),
    "topt_plan_create": (C.c_int, [C.POINTER(TProblem), _i32, _i32, C.POINTER(_vp)]),
    "topt_plan_destroy": (C.c_int, [_vp]),
    "topt_plan_set_graph": (C.c_int, [_vp, C.c_int32]),
    "topt_plan_advance": (C.c_int, [_vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _vp,
                                    C.POINTER(C.c_int32), C.POINTER(C.c_int64), _vp]),
    "topt_plan_filter": (C.c_int, [_vp, _i32, _vp]),
    "topt_plan_solve": (C.c_int, [_vp, _i32, _i32, C.POINTER(_i32), C.POINTER(_i32), _vp]),
    "topt_plan_sensitivity": (C.c_int, [_vp, _i32, _vp]),
    "topt_plan_oc": (C.c_int, [_vp, _vp]),
    "topt_plan_iteration": (C.c_int, [_vp, _i32, C.POINTER(TIterStats), _vp]),
    "topt_plan_stats": (C.c_int, [_vp, C.POINTER(TIterStats), _vp]),
    "topt_plan_set_timing": (C.c_int, [_vp, _i32]),
    "topt_plan_timing": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(_i64)]),
    "topt_plan_time_slot": (C.c_int, [_vp, _i32, _i32, C.POINTER(C.c_float), _vp]),
    "topt_op_workspace_bytes": (_i64, [_i32, _i32]),
    "topt_scale": (C.c_int, [C.POINTER(TGrid), _i32, _vp, _vp, _vp]),
    "topt_matvec": (C.c_int, [C.POINTER(TGrid), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt_diagonal": (C.c_int, [C.POINTER(TGrid), _i32, _vp, _vp, _vp, _vp]),
    "topt_block2_inverse": (C.c_int, [C.POINTER(TGrid), _i32, _vp, _vp, _vp, _vp, _vp]),
    "topt_elemblock_apply": (C.c_int, [C.POINTER(TGrid), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt_sensitivity": (C.c_int, [C.POINTER(TGrid), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt_filter": (C.c_int, [_i32, _i32, _d, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt_filter_tables": (C.c_int, [_i32, _i32, _d, _i32, _vp, _vp, _vp, _vp, _vp]),
    "topt_oc_update": (C.c_int, [_i32, _i32, _vp, _vp, _vp, _d, _d, _vp, _vp, _vp]),
    "topt_oc_update_noalloc": (C.c_int, [_i32, _i32, _vp, _vp, _vp, _d, _d, _vp, _vp]),
    "topt_volume": (C.c_int, [_i32, _i32, _vp, _vp, _vp, _vp]),
    "topt3_workspace_bytes": (_i64, [_i32, _i32, _i32]),
    "topt3_scale": (C.c_int, [C.POINTER(TGrid3), _vp, _vp, _vp]),
    "topt3_diagonal": (C.c_int, [C.POINTER(TGrid3), _vp, _vp, _vp, _vp]),
    "topt3_matvec": (C.c_int, [C.POINTER(TGrid3), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt3_elemblock_apply": (C.c_int, [C.POINTER(TGrid3), _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt3_sensitivity": (C.c_int, [C.POINTER(TGrid3), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt3_filter": (C.c_int, [_i32, _i32, _i32, _d, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "topt3_filter_tables": (C.c_int, [_i32, _i32, _i32, _d, _i32, _vp, _vp, _vp, _vp, _vp]),
    "topt3_solve": (C.c_int, [C.POINTER(TGrid3), _i32, _vp, _d, _vp, _vp, _vp, _vp, _i32, _d, _i32,
                              C.POINTER(_i32), C.POINTER(_i32), _vp, _vp]),
    "topt_blas1_workspace_bytes": (_i64, []),
    "topt_blas1": (C.c_int, [_i32, _i32, _i64, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}

B1 = {"fill": 0, "copy": 1, "axpy": 2, "scale": 3, "div": 4, "dot": 5, "maxdiff": 6, "sum": 7,
      "upd_jacobi": 8, "upd_none": 9, "upd_ur": 10, "upd_p": 11, "zero_fixed": 12}

_LIB: C.CDLL | None = None


def load() -> C.CDLL:
    """Load libtopt.so (built by ``__graft_entry__.build()``) and declare its ABI."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(hipcc, gfx950); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class ToptError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().topt_last_error().decode(errors="replace")
        raise ToptError(f"{what} failed ({rc}): {msg}")


def exported_symbols() -> list[str]:
    return list(_SIGS)
