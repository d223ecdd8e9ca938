"""Element stiffness matrices (same API as reference ``src/toptimization/material.py``).

Q4 plane stress on the unit square, 2x2 Gauss quadrature with
``D = E/(1-nu^2) [[1,nu,0],[nu,1,0],[0,0,(1-nu)/2]]`` (reference
``material.py:20-86``); H8 on the unit cube with 2x2x2 Gauss (``:89-171``).
DOF order per node is (x, y[, z]); node order as in ``mesh.Q4_CORNERS`` /
``mesh.H8_CORNERS``. The quadrature is written as one tensor contraction over
all Gauss points.
"""

from __future__ import annotations

import numpy as np

from .mesh import H8_CORNERS, Q4_CORNERS

_G = 1.0 / np.sqrt(3.0)


def _strain_matrices(corners: np.ndarray, points: np.ndarray) -> np.ndarray:
    """B matrices (n_pts, n_strain, n_dof) of the trilinear/bilinear unit element.

    Shape functions ``N_a = prod_k (1 + s_ak * xi_k) / 2^d`` with corner signs
    ``s = 2*corner - 1``; physical derivative = 2 * reference derivative on a
    unit element.
    """
    d = corners.shape[1]
    s = 2.0 * corners - 1.0                               # (n_nodes, d)
    fac = 1.0 + s[None, :, :] * points[:, None, :]         # (n_pts, n_nodes, d)
    grads = np.empty((points.shape[0], corners.shape[0], d))
    for k in range(d):
        others = np.prod(np.delete(fac, k, axis=2), axis=2)
        grads[:, :, k] = s[None, :, k] * others / 2.0 ** d * 2.0
    n_nodes = corners.shape[0]
    if d == 2:
        B = np.zeros((points.shape[0], 3, 2 * n_nodes))
        B[:, 0, 0::2] = grads[:, :, 0]
        B[:, 1, 1::2] = grads[:, :, 1]
        B[:, 2, 0::2] = grads[:, :, 1]
        B[:, 2, 1::2] = grads[:, :, 0]
        return B
    B = np.zeros((points.shape[0], 6, 3 * n_nodes))
    gx, gy, gz = grads[:, :, 0], grads[:, :, 1], grads[:, :, 2]
    B[:, 0, 0::3] = gx
    B[:, 1, 1::3] = gy
    B[:, 2, 2::3] = gz
    B[:, 3, 0::3], B[:, 3, 1::3] = gy, gx
    B[:, 4, 1::3], B[:, 4, 2::3] = gz, gy
    B[:, 5, 0::3], B[:, 5, 2::3] = gz, gx
    return B


def _gauss_points(d: int) -> np.ndarray:
    axes = np.meshgrid(*([np.array([-_G, _G])] * d), indexing="ij")
    return np.stack([a.ravel() for a in axes], axis=1)


def compute_Ke_2d(E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """8x8 float64 Q4 plane-stress stiffness (reference ``material.py:20-86``)."""
    D = E / (1.0 - nu * nu) * np.array([[1.0, nu, 0.0], [nu, 1.0, 0.0], [0.0, 0.0, 0.5 * (1.0 - nu)]])
    B = _strain_matrices(Q4_CORNERS.astype(float), _gauss_points(2))
    det_j = 0.25  # unit square mapped from [-1, 1]^2
    return det_j * np.einsum("pia,ij,pjb->ab", B, D, B)


def compute_Ke_3d(E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """24x24 float64 H8 stiffness (reference ``material.py:89-171``)."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2.0 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    B = _strain_matrices(H8_CORNERS.astype(float), _gauss_points(3))
    det_j = 0.125
    return det_j * np.einsum("pia,ij,pjb->ab", B, D, B)


def compute_Ke(E: float, nu: float, dim: int) -> np.ndarray:
    """Dispatch on dimension (reference ``material.py:174-179``)."""
    return compute_Ke_2d(E, nu) if dim == 2 else compute_Ke_3d(E, nu)
