"""Matrix-free PCG solve with the reference's signature (``fem/solver.py:27-159``).

``solve`` works on the module fields (``fields.u, f, rho, u_prev``) exactly as
the reference does: rho is the physical density, the result lands in
``fields.u`` and is copied to ``fields.u_prev``, the return value is the
reference's iteration count (``max_iter`` on non-convergence or breakdown,
0 for a zero load). The whole loop runs on the GPU engine: the device
computes alpha/beta/convergence, the host polls one status word per chunk.

Extra keyword ``precision``: "fp64" (default) or "mixed".
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from .. import ops
from ..engine import DesignState, Plan
from ..problem import PRECONDITIONERS, Problem
from . import fields as F

# device states of the current fields, keyed by the solver settings; all
# of them are dropped when fields.allocate() runs again, and at most
# _MAX_CACHED are kept (oldest evicted)
_cache: dict = {}
_cache_gen = -1
_MAX_CACHED = 4


def _drop(entry) -> None:
    entry[1].close()


def _state_for(E_min, penalty, max_iter, tol, preconditioner, precision):
    global _cache_gen
    if _cache_gen != F.GENERATION:
        for entry in _cache.values():
            _drop(entry)
        _cache.clear()
        _cache_gen = F.GENERATION
    key = (float(E_min), float(penalty), int(max_iter), float(tol), preconditioner, precision)
    hit = _cache.get(key)
    if hit is not None:
        return key, hit
    while len(_cache) >= _MAX_CACHED:
        _drop(_cache.pop(next(iter(_cache))))
    nelx, nely = F.grid()
    fixed = np.flatnonzero(F.is_fixed.cpu().numpy())
    f_np = F.f.double().cpu().numpy()
    nz = np.flatnonzero(f_np)
    prob = Problem(name="fields", dim=2, nelx=nelx, nely=nely, nelz=1, E=1.0, nu=0.3, E_min=float(E_min),
                   fixed_dofs=fixed, force_dofs=nz, force_values=f_np[nz], volfrac=0.5,
                   penalty=float(penalty), rmin=1.5, filter_type="density", max_iter=1, tol=0.0,
                   move_limit=0.2, max_cg_iter=int(max_iter), cg_tol=float(tol),
                   preconditioner=preconditioner, output_dir=Path("."), save_interval=1,
                   output_format="png", precision=precision)
    st = DesignState(prob, Ke=F.KE_NP, f=f_np)
    hit = _cache[key] = (st, Plan([st]))
    return key, hit


def solve(E_min: float, penalty: float, dim: int, max_iter: int = 2000, tol: float = 1e-8,
          preconditioner: str = "jacobi", warm_start: bool = True, *, precision: str = "fp64") -> int:
    """Solve K(F.rho) F.u = F.f; returns the number of CG iterations."""
    if preconditioner not in PRECONDITIONERS:
        raise ValueError(f"unknown preconditioner {preconditioner!r}")
    if dim != F.DIM:
        raise ValueError(f"dim={dim} does not match the allocated fields (dim {F.DIM})")
    if dim == 3:
        return _solve_3d(E_min, penalty, max_iter, tol, preconditioner, warm_start)
    key, (st, plan) = _state_for(E_min, penalty, max_iter, tol, preconditioner, precision)
    st.f.copy_(F.f)
    # |f|^2 sets the stopping tolerance and the zero-load test; the caller
    # may have changed f since the last call (solver.py:58-64 recomputes it
    # every time), and a plan holds the value it was created with
    fn2 = float(ops.blas1("dot", st.f.numel(), x=st.f, y=st.f, reduce=True)[0].item())
    if fn2 != st.desc.fnorm2:
        st.desc.fnorm2 = fn2
        plan.close()
        plan = Plan([st])
        _cache[key] = (st, plan)
    st.rhof.copy_(F.rho)
    if warm_start:
        st.u.copy_(F.u_prev)
    it, status = plan.solve(warm=1 if warm_start else 0, use_filtered=True)
    F.u.copy_(st.u)
    if status[0] != 4:            # zero load returns before u_prev is written (solver.py:60-63)
        F.u_prev.copy_(st.u)
    return int(it[0])


def _solve_3d(E_min, penalty, max_iter, tol, preconditioner, warm_start) -> int:
    """The 3-D solve: the whole PCG on the device (``h8.solve``), with the
    reference's field semantics (u from u_prev on a warm start, u_prev <- u
    unless the load is zero)."""
    import torch

    from .. import h8
    from .kernels import _fixmask, _grid3, _scale
    g = _grid3(F.Ke, E_min, penalty)
    s = _scale(g, F.rho, torch.float64)
    if warm_start:
        F.u.copy_(F.u_prev)
    n, status = h8.solve(g, preconditioner, F.f, _fixmask(F.is_fixed), s, F.u, warm=warm_start, tol=tol,
                         max_iter=max_iter)
    if status != 4:            # zero load returns before u_prev is written (solver.py:60-63)
        F.u_prev.copy_(F.u)
    return n
