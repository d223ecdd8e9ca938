"""ctypes access to oracle/liboracle.so (TEST INFRASTRUCTURE: the C
restatement of the reference hot path; see oracle/csrc/oracle_body.h)."""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "liboracle.so"
_L = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _L
    if _L is None:
        if not LIB.exists():
            from . import build
            build.build()
        L = C.CDLL(str(LIB))
        for suf, rp in (("f64", _dp), ("f32", _fp)):
            fn = getattr(L, f"orc_simp_{suf}")
            fn.restype = C.c_int
            fn.argtypes = [C.c_int, C.c_int] + [C.c_double] * 6 + [C.c_int, C.c_int, C.c_double, C.c_double,
                                                                     C.c_int, C.c_double, C.c_int, _up, _dp,
                                                                     rp, rp, _dp, _dp, _dp, _ip,
                                                                     C.POINTER(C.c_double)]
            mv = getattr(L, f"orc_matvec_{suf}")
            mv.restype = None
            mv.argtypes = [C.c_int, C.c_int, rp, rp, _up, rp, rp]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _L = L
    return _L


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def simp(prob, precision: str = "f64", max_iter: int | None = None, preconditioner: str | None = None):
    """Run the C SIMP loop for a Problem; returns dict like simp_np.simp_run
    plus 'seconds' (wall time of the loop)."""
    L = lib()
    real = np.float64 if precision == "f64" else np.float32
    fn = getattr(L, f"orc_simp_{precision}")
    nd, ne = prob.n_dofs, prob.n_elem
    fixed = np.zeros(nd, dtype=np.uint8)
    fixed[np.asarray(prob.fixed_dofs, dtype=np.int64)] = 1
    f = np.zeros(nd)
    np.add.at(f, np.asarray(prob.force_dofs, dtype=np.int64), np.asarray(prob.force_values, float))
    n_it = int(max_iter or prob.max_iter)
    pc = preconditioner or prob.preconditioner
    if pc not in ("jacobi", "none"):
        raise ValueError("the C oracle implements the jacobi and none preconditioners")
    rho = np.zeros(ne, dtype=real)
    u = np.zeros(nd, dtype=real)
    c = np.zeros(n_it)
    vol = np.zeros(n_it)
    ch = np.zeros(n_it)
    ncg = np.zeros(n_it, dtype=np.int32)
    secs = C.c_double()
    done = fn(prob.nelx, prob.nely, prob.E, prob.nu, prob.E_min, prob.penalty, prob.volfrac, prob.rmin,
              0 if prob.filter_type == "density" else 1, n_it, prob.tol, prob.move_limit, prob.max_cg_iter,
              prob.cg_tol, 1 if pc == "jacobi" else 0, fixed, f, rho, u, c, vol, ch, ncg, C.byref(secs))
    return {"c": c[:done].tolist(), "vol": vol[:done].tolist(), "change": ch[:done].tolist(),
            "n_cg": ncg[:done].tolist(), "rho": rho.astype(np.float64), "u": u.astype(np.float64),
            "seconds": secs.value, "iterations": done}
