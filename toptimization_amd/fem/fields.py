"""Device state with the reference's module-global field names.

Drop-in for the reference ``fem/fields.py:14-156``: ``allocate(...)`` with
the same arguments creates torch tensors on the GPU under the same global
names (``u, f, r, p, Ap, z, diag, u_prev, rho, rho_new, dc, dc_filt, dv,
rho_filt, Ke, Ke_inv, edof, is_fixed, diag_count, filt_neighbors,
filt_weights, filt_n_nb, _dot_result, _max_result``) and constants
(``N_DOFS, N_ELEM, DPE, KE_ALPHA_REG``). Fields are FP64 (the reference
uses FP32). The grid shape the HIP kernels need is read back from ``edof``
(``mesh.grid_from_edof`` for Q4, ``grid_from_edof_3d`` for H8); ``NELX,
NELY, NELZ, DIM`` and the fixed-DOF bitmask ``FIXMASK`` are set in addition.
"""

from __future__ import annotations

import numpy as np
import torch

from ..engine import fixed_bitmask, require_gpu
from ..mesh import grid_from_edof, grid_from_edof_3d

u = f = r = p = Ap = z = diag = u_prev = None
rho = rho_new = dc = dc_filt = dv = rho_filt = None
Ke = Ke_inv = edof = is_fixed = diag_count = None
filt_neighbors = filt_weights = filt_n_nb = None
_dot_result = _max_result = None
FIXMASK = None
N_DOFS = 0
N_ELEM = 0
DPE = 0
KE_ALPHA_REG = 0.0
NELX = 0
NELY = 0
NELZ = 1
DIM = 2
KE_NP = None
# incremented by every allocate(): device states built from the previous
# fields (fem/solver.py's cache) are discarded when it changes
GENERATION = 0


def allocate(n_dofs: int, n_elem: int, dpe: int, max_neighbors: int, Ke_np: np.ndarray,
             edof_np: np.ndarray, fixed_dofs: np.ndarray, force_dofs: np.ndarray,
             force_values: np.ndarray) -> None:
    """Allocate every field on the GPU and fill the static data (``fields.py:53-156``)."""
    global u, f, r, p, Ap, z, diag, u_prev, rho, rho_new, dc, dc_filt, dv, rho_filt
    global Ke, Ke_inv, edof, is_fixed, diag_count, filt_neighbors, filt_weights, filt_n_nb
    global _dot_result, _max_result, FIXMASK, N_DOFS, N_ELEM, DPE, KE_ALPHA_REG, NELX, NELY, KE_NP
    global GENERATION, NELZ, DIM
    if dpe not in (8, 24):
        raise ValueError(f"dpe must be 8 (Q4) or 24 (H8), got {dpe}")
    dev = require_gpu()
    N_DOFS, N_ELEM, DPE = int(n_dofs), int(n_elem), int(dpe)
    if dpe == 8:
        NELX, NELY = grid_from_edof(edof_np, N_ELEM)
        NELZ, DIM = 1, 2
    else:
        NELX, NELY, NELZ = grid_from_edof_3d(edof_np, N_ELEM)
        DIM = 3
    z64 = lambda n: torch.zeros(n, dtype=torch.float64, device=dev)  # noqa: E731
    u, f, r, p, Ap, z, diag, u_prev = (z64(N_DOFS) for _ in range(8))
    rho, rho_new, dc, dc_filt, dv, rho_filt = (z64(N_ELEM) for _ in range(6))
    KE_NP = np.asarray(Ke_np, dtype=np.float64)
    Ke = torch.as_tensor(KE_NP, device=dev)
    alpha = 0.1 * float(np.mean(np.abs(np.diag(KE_NP))))      # fields.py:139-143
    KE_ALPHA_REG = alpha
    Ke_inv = torch.as_tensor(np.linalg.inv(KE_NP + alpha * np.eye(dpe)), device=dev)
    edof = torch.as_tensor(np.asarray(edof_np, dtype=np.int32), device=dev)
    fixed = np.asarray(fixed_dofs, dtype=np.int64)
    fx = np.zeros(N_DOFS, dtype=np.int32)
    fx[fixed] = 1
    is_fixed = torch.as_tensor(fx, device=dev)
    FIXMASK = torch.as_tensor(fixed_bitmask(N_DOFS, fixed).view(np.int32), device=dev)
    diag_count = torch.zeros(N_DOFS, dtype=torch.int32, device=dev)
    filt_neighbors = torch.full((N_ELEM, max_neighbors), -1, dtype=torch.int32, device=dev)
    filt_weights = torch.zeros((N_ELEM, max_neighbors), dtype=torch.float32, device=dev)
    filt_n_nb = torch.zeros(N_ELEM, dtype=torch.int32, device=dev)
    _dot_result = z64(2)
    _max_result = z64(1)
    f_np = np.zeros(N_DOFS)
    np.add.at(f_np, np.asarray(force_dofs, dtype=np.int64), np.asarray(force_values, dtype=np.float64))
    f_np[fixed] = 0.0
    f.copy_(torch.from_numpy(f_np))
    GENERATION += 1


def grid() -> tuple[int, int]:
    """(nelx, nely) of the allocated 2-D grid."""
    if NELX == 0:
        raise RuntimeError("fields.allocate() has not been called")
    if DIM != 2:
        raise RuntimeError("the allocated fields are 3-D (use grid3())")
    return NELX, NELY


def grid3() -> tuple[int, int, int]:
    """(nelx, nely, nelz) of the allocated 3-D grid."""
    if NELX == 0:
        raise RuntimeError("fields.allocate() has not been called")
    if DIM != 3:
        raise RuntimeError("the allocated fields are 2-D (use grid())")
    return NELX, NELY, NELZ


def elem_grid() -> tuple[int, int]:
    """Element array as (columns, rows) for the element-only kernels (OC,
    volume): (nelx, nely) in 2-D, (nelx, nely*nelz) in 3-D."""
    if NELX == 0:
        raise RuntimeError("fields.allocate() has not been called")
    return NELX, NELY * NELZ
