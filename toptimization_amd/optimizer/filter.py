"""Density / sensitivity filter with the reference's API (``optimizer/filter.py``).

The HIP kernels use the compact stencil (offsets + FP32 weights in the
reference's table order); ``precompute_filter`` still fills the reference's
padded tables (``fields.filt_neighbors/weights/n_nb``) by expanding that
stencil on the GPU, bit-identical to the reference's Python loop.
"""

from __future__ import annotations

from .. import _lib, h8, ops
from ..fem import fields as F

RMIN: float | None = None


_G3: dict = {}


def _grid3():
    """Descriptor and workspace of the 3-D grid (the filter uses no Ke)."""
    key = (F.GENERATION, *F.grid3())
    if _G3.get("key") != key:
        _G3["key"], _G3["g"] = key, h8.Grid3(*key[1:])
    return _G3["g"]


def precompute_filter(problem, centers) -> int:
    """Neighbour tables of ``filter.py:22-103``; returns the actual max count."""
    global RMIN
    RMIN = float(problem.rmin)
    if problem.dim == 3:
        nb, wt, cnt = h8.filter_tables(_grid3(), RMIN)
    else:
        nb, wt, cnt = ops.filter_tables(problem.nelx, problem.nely, RMIN)
    width = F.filt_neighbors.shape[1]
    F.filt_neighbors.copy_(nb[:, :width])
    F.filt_weights.copy_(wt[:, :width])
    F.filt_n_nb.copy_(cnt)
    return int(cnt.max().item())


def _rmin() -> float:
    if RMIN is None:
        raise RuntimeError("precompute_filter() has not been called")
    return RMIN


def apply_density_filter(rho, rho_filt, neighbors, weights, n_nb, n_elem: int) -> None:
    """rho_filt = H rho / (Hs + 1e-20) (``filter.py:106-124``)."""
    if F.DIM == 3:
        rho_filt.copy_(h8.filter_apply(_grid3(), _rmin(), _lib.FILTER_DENSITY, rho=rho.double().contiguous()))
        return
    nelx, nely = F.grid()
    rho_filt.copy_(ops.filter_apply(nelx, nely, _rmin(), _lib.FILTER_DENSITY, rho=rho.double().contiguous()))


def apply_dc_average_filter(dc_in, dc_out, neighbors, weights, n_nb, n_elem: int) -> None:
    """dc_out = H dc / (Hs + 1e-20) (``filter.py:127-149``)."""
    if F.DIM == 3:
        dc_out.copy_(h8.filter_apply(_grid3(), _rmin(), _lib.FILTER_DCAVG, dc=dc_in.double().contiguous()))
        return
    nelx, nely = F.grid()
    dc_out.copy_(ops.filter_apply(nelx, nely, _rmin(), _lib.FILTER_DCAVG, dc=dc_in.double().contiguous()))


def apply_sensitivity_filter(rho, dc, dc_filt, neighbors, weights, n_nb, n_elem: int) -> None:
    """dc_filt = H (rho dc) / (rho Hs + 1e-20) (``filter.py:152-171``)."""
    if F.DIM == 3:
        dc_filt.copy_(h8.filter_apply(_grid3(), _rmin(), _lib.FILTER_SENS, rho=rho.double().contiguous(),
                                      dc=dc.double().contiguous()))
        return
    nelx, nely = F.grid()
    dc_filt.copy_(ops.filter_apply(nelx, nely, _rmin(), _lib.FILTER_SENS, rho=rho.double().contiguous(),
                                   dc=dc.double().contiguous()))


def apply_filter_transpose(dc_in, dc_out) -> None:
    """Exact chain rule of the density filter, H^T (dc / (Hs + 1e-20)) (no reference twin)."""
    nelx, nely = F.grid()
    dc_out.copy_(ops.filter_apply(nelx, nely, _rmin(), _lib.FILTER_TRANSPOSE, dc=dc_in.double().contiguous()))


def apply_dc_filter_density_mode(n_elem: int) -> None:
    """``filter.py:174-180``."""
    apply_dc_average_filter(F.dc, F.dc_filt, F.filt_neighbors, F.filt_weights, F.filt_n_nb, n_elem)


def apply_dc_filter_sensitivity_mode(n_elem: int) -> None:
    """``filter.py:183-189``."""
    apply_sensitivity_filter(F.rho, F.dc, F.dc_filt, F.filt_neighbors, F.filt_weights, F.filt_n_nb, n_elem)
