"""Matrix-free FEM kernels with the reference's call signatures.

Drop-in for ``fem/kernels.py`` of the reference: every function takes the
same arguments (torch CUDA tensors where the reference takes Taichi fields)
and runs a HIP kernel of libtopt. Scalar-returning functions return Python
floats (a device->host read, as in the reference). ``dim == 2`` runs the Q4
kernels, ``dim == 3`` the H8 ones (``h8.py``).
"""

from __future__ import annotations

import numpy as np
import torch

from .. import h8, ops
from . import fields as F


def _dim2(dim: int) -> None:
    if dim != 2:
        raise ValueError(f"dim must be 2 here, got {dim}")


def _check_dim(dim: int) -> int:
    if dim not in (2, 3) or dim != F.DIM:
        raise ValueError(f"dim={dim} does not match the allocated fields (dim {F.DIM})")
    return dim


_grid3_cache: dict = {}


def _grid3(Ke, E_min: float, penalty: float):
    """h8.Grid3 for the current 3-D fields (Ke on the device, workspace)."""
    nelx, nely, nelz = F.grid3()
    Ke_np = Ke.detach().cpu().numpy() if isinstance(Ke, torch.Tensor) else np.asarray(Ke)
    key = (F.GENERATION, nelx, nely, nelz, float(E_min), float(penalty), Ke_np.tobytes())
    g = _grid3_cache.get(key)
    if g is None:
        _grid3_cache.clear()
        g = _grid3_cache[key] = h8.Grid3(nelx, nely, nelz, Ke_np.reshape(24, 24), E_min, penalty)
    return g


_grid_cache: dict = {}


def _grid(Ke, E_min: float, penalty: float, edof=None):
    """topt_grid for the current fields (grid shape from fields.allocate)."""
    nelx, nely = F.grid()
    Ke_np = Ke.detach().cpu().numpy() if isinstance(Ke, torch.Tensor) else np.asarray(Ke)
    key = (nelx, nely, float(E_min), float(penalty), Ke_np.tobytes())
    g = _grid_cache.get(key)
    if g is None:
        g = _grid_cache[key] = ops.make_grid(nelx, nely, Ke_np, E_min, penalty)
    return g


_s_bufs: dict = {}


def _scale(g, rho, dtype):
    """SIMP scales of rho in a buffer reused across calls (one launch, no
    allocation after the first call for a given length and dtype)."""
    key = (rho.numel(), dtype, rho.device)
    buf = _s_bufs.get(key)
    if buf is None:
        buf = _s_bufs[key] = torch.empty(rho.numel(), dtype=dtype, device=rho.device)
    if isinstance(g, h8.Grid3):
        return h8.scale(g, rho.double().contiguous(), out=buf)
    return ops.scale(g, rho, dtype, out=buf)


def _fixmask(is_fixed):
    if is_fixed is F.is_fixed and F.FIXMASK is not None:
        return F.FIXMASK
    return ops.fix_mask(is_fixed.numel(), np.flatnonzero(is_fixed.cpu().numpy()), is_fixed.device)


def fill_scalar(field, val: float) -> None:
    """field[:] = val (``kernels.py:19-23``)."""
    ops.blas1("fill", field.numel(), val, x=field)


def copy_field(src, dst) -> None:
    """dst = src (``kernels.py:26-30``)."""
    ops.blas1("copy", src.numel(), x=src, y=dst)


def axpy(a: float, x, y, result) -> None:
    """result = a x + y (``kernels.py:33-37``)."""
    ops.blas1("axpy", result.numel(), a, x=x, y=y, z=result)


def scale_field(a: float, x, result) -> None:
    """result = a x (``kernels.py:40-44``)."""
    ops.blas1("scale", result.numel(), a, x=x, z=result)


def dot_product(a, b) -> float:
    """sum a_i b_i, FP64 deterministic (``kernels.py:47-53``)."""
    return float(ops.blas1("dot", a.numel(), x=a, y=b, reduce=True)[0].item())


def l2_norm_sq(a) -> float:
    """sum a_i^2 (``kernels.py:56-62``)."""
    return float(ops.blas1("dot", a.numel(), x=a, y=a, reduce=True)[0].item())


def apply_diag_precond(diag, r, z) -> None:
    """z = r / diag (``kernels.py:65-73``)."""
    ops.blas1("div", z.numel(), x=r, y=diag, z=z)


def compute_diagonal(diag, rho, Ke, edof, is_fixed, E_min: float, penalty: float, n_elem: int,
                     dim: int) -> None:
    """diag K(rho), fixed -> 1, clamp 1e-20 (``kernels.py:76-140``)."""
    if _check_dim(dim) == 3:
        g = _grid3(Ke, E_min, penalty)
        diag.copy_(h8.diagonal(g, _scale(g, rho, torch.float64), _fixmask(is_fixed)))
        return
    g = _grid(Ke, E_min, penalty)
    s = _scale(g, rho, diag.dtype)
    diag.copy_(ops.diagonal(g, s, _fixmask(is_fixed)))


def compute_matvec(p_in, Ap_out, rho, Ke, edof, is_fixed, E_min: float, penalty: float,
                   n_elem: int, dim: int) -> None:
    """Ap = K(rho) p with identity rows at fixed DOFs (``kernels.py:143-231``)."""
    if _check_dim(dim) == 3:
        g = _grid3(Ke, E_min, penalty)
        h8.matvec(g, p_in, _scale(g, rho, torch.float64), _fixmask(is_fixed), Ap=Ap_out)
        return
    g = _grid(Ke, E_min, penalty)
    s = _scale(g, rho, p_in.dtype)
    ops.matvec(g, p_in, s, _fixmask(is_fixed), Ap=Ap_out)


def compute_sensitivity(u, rho, dc, dv, Ke, edof, E_min: float, penalty: float, n_elem: int,
                        dim: int) -> float:
    """dc, dv = 1 and the compliance (``kernels.py:234-327``)."""
    if _check_dim(dim) == 3:
        g = _grid3(Ke, E_min, penalty)
        c, dcv = h8.sensitivity(g, u.double().contiguous(), rho.double().contiguous())
        dc.copy_(dcv)
        dv.fill_(1.0)
        return float(c.item())
    g = _grid(Ke, E_min, penalty)
    c, dcv = ops.sensitivity(g, u.double().contiguous(), rho.double().contiguous())
    dc.copy_(dcv)
    dv.fill_(1.0)
    return float(c.item())


def compute_volume(rho, n_elem: int) -> float:
    """mean rho (``kernels.py:330-336``)."""
    return float(ops.blas1("sum", n_elem, x=rho, reduce=True)[0].item()) / n_elem


def max_density_change(rho, rho_new, n_elem: int) -> float:
    """max |rho_new - rho| (``kernels.py:339-345``)."""
    return float(ops.blas1("maxdiff", n_elem, x=rho_new, y=rho, reduce=True)[0].item())
