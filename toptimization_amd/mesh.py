"""Structured-grid connectivity (same API as the reference ``src/toptimization/mesh.py``).

2-D layout (reference ``mesh.py:3-24``): node ``(ix, iy)`` -> ``ix*(nely+1)+iy``,
element ``(ex, ey)`` -> ``ex*nely+ey``, local element DOFs
``[BLx, BLy, BRx, BRy, TRx, TRy, TLx, TLy]``.

The GPU kernels never read ``edof``: they derive every neighbour from the grid
indices. ``build_edof`` exists for drop-in callers and for the tests that pin
the numbering (reference ``tests/test_mesh.py``).
"""

from __future__ import annotations

import numpy as np

from .problem import Problem

# corner offsets (dx, dy) of the four element nodes in local order BL, BR, TR, TL
Q4_CORNERS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.int64)
# 3-D H8 corners: bottom face CCW then top face CCW (reference mesh.py:86-95)
H8_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                       [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)


def build_edof_2d(problem: Problem) -> np.ndarray:
    """(n_elem, 8) int32 element -> global DOF map (reference ``mesh.py:33-65``)."""
    nelx, nely = problem.nelx, problem.nely
    ex, ey = np.divmod(np.arange(nelx * nely, dtype=np.int64), nely)
    nodes = (ex[:, None] + Q4_CORNERS[None, :, 0]) * (nely + 1) + (ey[:, None] + Q4_CORNERS[None, :, 1])
    dofs = 2 * nodes[:, :, None] + np.arange(2)[None, None, :]
    return dofs.reshape(-1, 8).astype(np.int32)


def build_edof_3d(problem: Problem) -> np.ndarray:
    """(n_elem, 24) int32 H8 connectivity (reference ``mesh.py:68-112``)."""
    nelx, nely, nelz = problem.nelx, problem.nely, problem.nelz
    e = np.arange(nelx * nely * nelz, dtype=np.int64)
    ex, rem = np.divmod(e, nely * nelz)
    ey, ez = np.divmod(rem, nelz)
    cx = ex[:, None] + H8_CORNERS[None, :, 0]
    cy = ey[:, None] + H8_CORNERS[None, :, 1]
    cz = ez[:, None] + H8_CORNERS[None, :, 2]
    nodes = (cx * (nely + 1) + cy) * (nelz + 1) + cz
    dofs = 3 * nodes[:, :, None] + np.arange(3)[None, None, :]
    return dofs.reshape(-1, 24).astype(np.int32)


def build_edof(problem: Problem) -> np.ndarray:
    """Connectivity dispatched on dimension (reference ``mesh.py:115-120``)."""
    return build_edof_2d(problem) if problem.dim == 2 else build_edof_3d(problem)


def element_centers_2d(problem: Problem) -> np.ndarray:
    """(n_elem, 2) float32 element centres ``(ex+0.5, ey+0.5)`` (``mesh.py:123-136``)."""
    ex, ey = np.divmod(np.arange(problem.nelx * problem.nely), problem.nely)
    return (np.stack([ex, ey], axis=1) + 0.5).astype(np.float32)


def element_centers_3d(problem: Problem) -> np.ndarray:
    """(n_elem, 3) float32 centres (``mesh.py:139-154``)."""
    n = problem.nelx * problem.nely * problem.nelz
    ex, rem = np.divmod(np.arange(n), problem.nely * problem.nelz)
    ey, ez = np.divmod(rem, problem.nelz)
    return (np.stack([ex, ey, ez], axis=1) + 0.5).astype(np.float32)


def element_centers(problem: Problem) -> np.ndarray:
    return element_centers_2d(problem) if problem.dim == 2 else element_centers_3d(problem)


def grid_from_edof(edof: np.ndarray, n_elem: int) -> tuple[int, int]:
    """Recover (nelx, nely) from a 2-D structured ``edof`` and check it is one.

    The reference ``fields.allocate`` (``fem/fields.py:53-63``) receives only
    ``edof``; the grid kernels need the grid shape, so it is read back from the
    first element (its BR x-DOF is ``2*(nely+1)``) and the whole map is
    verified against the structured numbering.
    """
    edof = np.asarray(edof)
    if edof.ndim != 2 or edof.shape != (n_elem, 8):
        raise ValueError(f"expected a (n_elem, 8) Q4 edof, got {edof.shape}")
    nely = int(edof[0, 2]) // 2 - 1
    if nely <= 0 or n_elem % nely:
        raise ValueError("edof is not a structured Q4 grid numbering")
    nelx = n_elem // nely
    probe = Problem.__new__(Problem)
    probe.nelx, probe.nely = nelx, nely
    if not np.array_equal(build_edof_2d(probe), edof.astype(np.int32)):
        raise ValueError("edof is not a structured Q4 grid numbering")
    return nelx, nely


def grid_from_edof_3d(edof: np.ndarray, n_elem: int) -> tuple[int, int, int]:
    """Recover (nelx, nely, nelz) from a 3-D structured ``edof`` and check it is
    one: element 0's local node 3 is (0, 1, 0), node ``nelz+1``, and its node
    1 is (1, 0, 0), node ``(nely+1)(nelz+1)`` (``mesh.py:91-100``)."""
    edof = np.asarray(edof)
    if edof.ndim != 2 or edof.shape != (n_elem, 24):
        raise ValueError(f"expected a (n_elem, 24) H8 edof, got {edof.shape}")
    nelz = int(edof[0, 9]) // 3 - 1
    if nelz <= 0:
        raise ValueError("edof is not a structured H8 grid numbering")
    nely = int(edof[0, 3]) // (3 * (nelz + 1)) - 1
    if nely <= 0 or n_elem % (nely * nelz):
        raise ValueError("edof is not a structured H8 grid numbering")
    nelx = n_elem // (nely * nelz)
    probe = Problem.__new__(Problem)
    probe.nelx, probe.nely, probe.nelz = nelx, nely, nelz
    if not np.array_equal(build_edof_3d(probe), edof.astype(np.int32)):
        raise ValueError("edof is not a structured H8 grid numbering")
    return nelx, nely, nelz
