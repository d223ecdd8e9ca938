"""Element-block Jacobi preconditioner with the reference's call signatures
(``fem/preconditioner.py``). The element multiplicities are closed-form on
the structured grid; ``setup`` fills ``diag_count`` for drop-in callers."""

from __future__ import annotations

import numpy as np
import torch

from .. import ops
from . import fields as F
from .kernels import _dim2, _fixmask, _grid, _scale


def setup(edof, diag_count, n_elem: int, dim: int) -> None:
    """Elements sharing each DOF: 4 interior, 2 edge, 1 corner (``preconditioner.py:159-164``)."""
    _dim2(dim)
    nelx, nely = F.grid()
    ix, iy = np.divmod(np.arange((nelx + 1) * (nely + 1)), nely + 1)
    cx = np.where((ix == 0) | (ix == nelx), 1, 2)
    cy = np.where((iy == 0) | (iy == nely), 1, 2)
    diag_count.copy_(torch.as_tensor(np.repeat(cx * cy, 2).astype(np.int32)))


def apply_and_rz(r, z, rho, Ke_inv, edof, diag_count, is_fixed, dot_result, E_min: float,
                 penalty: float, alpha_reg: float, n_elem: int, dim: int) -> float:
    """z = M r (symmetric element-block Jacobi), returns r.z and writes |r|^2
    to dot_result[0] (``preconditioner.py:47-108, 167-181``). The regularised
    Ke inverse is formed from fields.Ke inside libtopt (``fields.py:139-143``)."""
    _dim2(dim)
    g = _grid(F.Ke, E_min, penalty)
    s = _scale(g, rho, r.dtype)
    zz, rz, rr = ops.elemblock_apply(g, r, s, _fixmask(is_fixed))
    z.copy_(zz)
    dot_result.view(-1)[0] = rr
    return float(rz.item())
