"""Optimality-criteria update with the reference's API (``optimizer/oc.py:52-93``)."""

from __future__ import annotations

from .. import ops
from ..fem import fields as F


def oc_update(problem) -> float:
    """Bisection on lambda over ``fields.dc_filt`` (dv = 1); updates
    ``fields.rho`` in place and returns the max density change."""
    nelx, nely = F.elem_grid()       # 3-D: the OC update only reads element arrays
    rho = F.rho
    dcf = F.dc_filt.double().contiguous()
    change, vol, lam, steps = ops.oc_update(nelx, nely, rho, dcf, problem.volfrac, problem.move_limit)
    F.rho_new.copy_(rho)
    return change
