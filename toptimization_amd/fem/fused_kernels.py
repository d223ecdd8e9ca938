"""Fused PCG kernels with the reference's call signatures
(``fem/fused_kernels.py``); each is one libtopt launch."""

from __future__ import annotations

import torch

from .. import h8, ops
from . import fields as F
from .kernels import _check_dim, _fixmask, _grid, _grid3, _scale


def matvec_and_dot(p_in, Ap_out, rho, Ke, edof, is_fixed, E_min: float, penalty: float,
                   n_elem: int, dim: int) -> float:
    """Ap = K(rho) p; returns p.Ap (``fused_kernels.py:21-94``)."""
    if _check_dim(dim) == 3:
        g = _grid3(Ke, E_min, penalty)
        _, dot = h8.matvec(g, p_in, _scale(g, rho, torch.float64), _fixmask(is_fixed), Ap=Ap_out, with_dot=True)
        return float(dot.item())
    g = _grid(Ke, E_min, penalty)
    s = _scale(g, rho, p_in.dtype)
    _, dot = ops.matvec(g, p_in, s, _fixmask(is_fixed), Ap=Ap_out, with_dot=True)
    return float(dot.item())


def update_and_precond_jacobi(u, r, z, p, Ap, diag, dot_result, alpha: float) -> float:
    """u += a p; r -= a Ap; z = r / diag; returns r.z, writes |r|^2 to
    dot_result[0] (``fused_kernels.py:110-139``)."""
    red = ops.blas1("upd_jacobi", u.numel(), alpha, x=u, y=r, z=z, w=p, v=Ap, d=diag, reduce=True)
    dot_result.view(-1)[0] = red[1]
    return float(red[0].item())


def update_and_precond_none(u, r, z, p, Ap, dot_result, alpha: float) -> float:
    """As above with z = r (``fused_kernels.py:142-166``)."""
    red = ops.blas1("upd_none", u.numel(), alpha, x=u, y=r, z=z, w=p, v=Ap, reduce=True)
    dot_result.view(-1)[0] = red[1]
    return float(red[0].item())


def update_ur(u, r, p, Ap, alpha: float) -> None:
    """u += a p; r -= a Ap (``fused_kernels.py:169-180``)."""
    ops.blas1("upd_ur", u.numel(), alpha, x=u, y=r, w=p, v=Ap)


def update_p(z, p, beta: float) -> None:
    """p = z + beta p (``fused_kernels.py:183-187``)."""
    ops.blas1("upd_p", p.numel(), beta, z=z, w=p)
