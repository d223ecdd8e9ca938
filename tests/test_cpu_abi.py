"""CPU-side checks of the C ABI: the library loads, exports every entry point
that include/topt.h declares, and the layout queries work without a GPU."""

import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "topt.h").read_text()
    return sorted(set(re.findall(r"\b(topt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from toptimization_amd import _lib
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.exported_symbols()) | {"topt_last_error"}


def test_layout_and_workspace_queries():
    from toptimization_amd import _lib
    lib = _lib.load()
    offs = (C.c_int64 * len(_lib.BUF))()
    total = C.c_int64()
    for dtype in (_lib.F32, _lib.F64):
        for pc in (0, 1, 2, 3):
            assert lib.topt_problem_layout(1200, 400, dtype, pc, 6.0, offs, C.byref(total)) == 0
            o = list(offs)
            assert all(v % 256 == 0 for v in o)
            assert total.value > 8 * 2 * 1201 * 401 * 6
    assert lib.topt_problem_layout(0, 4, 1, 1, 1.5, offs, C.byref(total)) != 0
    assert b"nelx" in lib.topt_last_error()
    assert lib.topt_problem_layout(4, 4, 7, 1, 1.5, offs, C.byref(total)) != 0
    assert lib.topt_op_workspace_bytes(60, 20) > 0
    assert lib.topt_op_workspace_bytes(0, 20) < 0


def test_gpu_entry_points_reject_null_without_touching_the_gpu():
    from toptimization_amd import _lib
    lib = _lib.load()
    assert lib.topt_plan_create(None, 0, 1, None) == _lib.load().topt_plan_create(None, 0, 1, None)
    assert lib.topt_plan_create(None, 0, 1, None) < 0
    assert lib.topt_matvec(None, 1, None, None, None, None, None, None, None) < 0
    assert lib.topt_oc_update(4, 4, None, None, None, 0.5, 0.2, None, None, None) < 0


def test_noalloc_oc_rejects_oversize_grid_on_the_host():
    """The size check runs before any device call (no GPU needed)."""
    from toptimization_amd import _lib
    lib = _lib.load()
    dummy = C.c_void_p(256)
    # 8192 x 8192 elements need ~8.5 MB of scratch, more than the 4 MiB array
    assert lib.topt_oc_update_noalloc(8192, 8192, dummy, dummy, None, 0.5, 0.2, dummy, None) == -1
    assert b"static" in lib.topt_last_error()
    assert lib.topt_oc_update_noalloc(4, 4, None, None, None, 0.5, 0.2, None, None) == -1
    assert lib.topt_oc_update_noalloc(4, 4, dummy, dummy, dummy, 0.5, 0.2, dummy, None) == -1
    assert b"dv must be NULL" in lib.topt_last_error()
