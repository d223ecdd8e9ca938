"""Element-block Jacobi preconditioner with the reference's call signatures
(``fem/preconditioner.py``). The element multiplicities are closed-form on
the structured grid; ``setup`` fills ``diag_count`` for drop-in callers."""

from __future__ import annotations

import numpy as np
import torch

from .. import h8, ops
from . import fields as F
from .kernels import _check_dim, _fixmask, _grid, _grid3, _scale


def setup(edof, diag_count, n_elem: int, dim: int) -> None:
    """Elements sharing each DOF (``preconditioner.py:19-44, 159-164``): the
    product over the axes of 1 at the grid's ends and 2 inside."""
    if _check_dim(dim) == 3:
        nelx, nely, nelz = F.grid3()
        n = np.arange((nelx + 1) * (nely + 1) * (nelz + 1))
        iz = n % (nelz + 1)
        ix, iy = np.divmod(n // (nelz + 1), nely + 1)
        per = [np.where((i == 0) | (i == m), 1, 2) for i, m in ((ix, nelx), (iy, nely), (iz, nelz))]
        diag_count.copy_(torch.as_tensor(np.repeat(per[0] * per[1] * per[2], 3).astype(np.int32)))
        return
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
    if _check_dim(dim) == 3:
        g = _grid3(F.Ke, E_min, penalty)
        zz, rz, rr = h8.elemblock_apply(g, r, _scale(g, rho, torch.float64), _fixmask(is_fixed))
        z.copy_(zz)
        dot_result.view(-1)[0] = rr
        return float(rz.item())
    g = _grid(F.Ke, E_min, penalty)
    s = _scale(g, rho, r.dtype)
    zz, rz, rr = ops.elemblock_apply(g, r, s, _fixmask(is_fixed))
    z.copy_(zz)
    dot_result.view(-1)[0] = rr
    return float(rz.item())
