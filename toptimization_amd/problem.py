"""Problem definition: YAML problem files -> boundary-condition and load arrays.

Same public surface as the reference parser (``src/toptimization/problem.py``):
``Problem`` (dataclass, ``problem.py:18-78``) and ``load_problem(path)``
(``problem.py:81-164``), with the same defaults, validation errors and DOF
numbering (node ``(ix, iy)`` -> ``ix*(nely+1)+iy``, DOF ``dim*node + comp``).

Additions for the GPU engine (documented, all optional in YAML):

* ``solver.preconditioner`` also accepts ``block_jacobi_2x2`` (nodal 2x2 block
  Jacobi) and ``multigrid`` (geometric multigrid V-cycle with Galerkin coarse
  levels); neither has a reference counterpart.
* ``solver.precision``: ``fp64`` (plain FP64 PCG, default) or ``mixed``
  (FP32 matvec/vectors, FP64 dots and FP64 true-residual refinement).
* ``make_problem(...)`` builds a ``Problem`` programmatically (serving requests).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Literal

import numpy as np
import yaml

PRECONDITIONERS = ("jacobi", "block_jacobi", "none", "block_jacobi_2x2", "multigrid")
PRECISIONS = ("fp64", "mixed")


@dataclass
class Problem:
    """All numerical parameters of one compliance topology-optimization run.

    Field names and meaning follow ``problem.py:18-78`` of the reference.
    """

    name: str
    dim: Literal[2, 3]
    nelx: int
    nely: int
    nelz: int
    E: float
    nu: float
    E_min: float
    fixed_dofs: np.ndarray
    force_dofs: np.ndarray
    force_values: np.ndarray
    volfrac: float
    penalty: float
    rmin: float
    filter_type: Literal["density", "sensitivity"]
    max_iter: int
    tol: float
    move_limit: float
    max_cg_iter: int
    cg_tol: float
    preconditioner: str
    output_dir: Path
    save_interval: int
    output_format: Literal["png", "vtk"]
    precision: str = "fp64"

    n_nodes: int = field(init=False)
    n_dofs: int = field(init=False)
    n_elem: int = field(init=False)
    free_dofs: np.ndarray = field(init=False)
    dofs_per_node: int = field(init=False)

    def __post_init__(self) -> None:
        self.dofs_per_node = int(self.dim)
        shape_nodes = [self.nelx + 1, self.nely + 1]
        shape_elems = [self.nelx, self.nely]
        if self.dim == 3:
            shape_nodes.append(self.nelz + 1)
            shape_elems.append(self.nelz)
        self.n_nodes = int(np.prod(shape_nodes))
        self.n_elem = int(np.prod(shape_elems))
        self.n_dofs = self.dofs_per_node * self.n_nodes
        fixed_mask = np.zeros(self.n_dofs, dtype=bool)
        fixed_mask[np.asarray(self.fixed_dofs, dtype=np.int64)] = True
        self.free_dofs = np.flatnonzero(~fixed_mask)


# --------------------------------------------------------------------------
# node numbering helpers
# --------------------------------------------------------------------------

def node_id_2d(ix, iy, nely: int):
    """Global node id of grid node (ix, iy): column-major, ``problem.py:167-173``."""
    return ix * (nely + 1) + iy


def node_id_3d(ix, iy, iz, nely: int, nelz: int):
    """Global node id of 3-D grid node, ``problem.py:314-315``."""
    return (ix * (nely + 1) + iy) * (nelz + 1) + iz


_EDGES_2D = {
    # edge name -> (ix range, iy range) as functions of (nelx, nely)
    "left": lambda nx, ny: (np.array([0]), np.arange(ny + 1)),
    "right": lambda nx, ny: (np.array([nx]), np.arange(ny + 1)),
    "bottom": lambda nx, ny: (np.arange(nx + 1), np.array([0])),
    "top": lambda nx, ny: (np.arange(nx + 1), np.array([ny])),
}

_FACES_3D = {
    # face name -> fixed axis and its value selector
    "left": (0, lambda n: 0), "right": (0, lambda n: n),
    "bottom": (1, lambda n: 0), "top": (1, lambda n: n),
    "front": (2, lambda n: 0), "back": (2, lambda n: n),
}

_CORNERS_2D = {
    "bottom-left": lambda nx, ny: (0, 0),
    "bottom-right": lambda nx, ny: (nx, 0),
    "top-left": lambda nx, ny: (0, ny),
    "top-right": lambda nx, ny: (nx, ny),
}

_CORNERS_3D = {
    "front-bottom-left": lambda nx, ny, nz: (0, 0, 0),
    "front-bottom-right": lambda nx, ny, nz: (nx, 0, 0),
    "back-bottom-left": lambda nx, ny, nz: (0, 0, nz),
    "back-bottom-right": lambda nx, ny, nz: (nx, 0, nz),
}


def _grid_ids_2d(ix_range, iy_range, nely):
    """Node ids for the tensor product ix_range x iy_range, ix-major order."""
    ix, iy = np.meshgrid(ix_range, iy_range, indexing="ij")
    return [int(v) for v in node_id_2d(ix, iy, nely).ravel()]


def _grid_ids_3d(ranges, nely, nelz):
    ix, iy, iz = np.meshgrid(*ranges, indexing="ij")
    return [int(v) for v in node_id_3d(ix, iy, iz, nely, nelz).ravel()]


def resolve_nodes(spec: dict, nelx: int, nely: int, nelz: int, dim: int) -> list[int]:
    """Node ids selected by one YAML node spec (``problem.py:244-263``).

    Supported: ``edge`` (2-D), ``face`` (3-D), ``point`` (``location`` clamped
    to the grid in 2-D, or named ``corner``), ``region`` (inclusive box).
    """
    kind = spec["type"]
    if kind == "edge" and dim == 2:
        name = spec["edge"]
        if name not in _EDGES_2D:
            raise ValueError(f"Unknown edge '{name}'. Use 'left', 'right', 'top', 'bottom'.")
        return _grid_ids_2d(*_EDGES_2D[name](nelx, nely), nely)
    if kind == "face" and dim == 3:
        name = spec["face"]
        if name not in _FACES_3D:
            raise ValueError(f"Unknown face '{name}'.")
        axis, pick = _FACES_3D[name]
        sizes = (nelx, nely, nelz)
        ranges = [np.arange(s + 1) for s in sizes]
        ranges[axis] = np.array([pick(sizes[axis])])
        return _grid_ids_3d(ranges, nely, nelz)
    if kind == "point":
        return [_point_node(spec, nelx, nely, nelz, dim)]
    if kind == "region":
        lo_hi = [spec.get("x", [0, nelx]), spec.get("y", [0, nely])]
        if dim == 2:
            rx, ry = (np.arange(int(a), int(b) + 1) for a, b in lo_hi)
            return _grid_ids_2d(rx, ry, nely)
        lo_hi.append(spec.get("z", [0, nelz]))
        return _grid_ids_3d([np.arange(int(a), int(b) + 1) for a, b in lo_hi], nely, nelz)
    raise ValueError(f"Unknown node spec type '{kind}' for dim={dim}")


def _point_node(spec, nelx, nely, nelz, dim) -> int:
    if "location" in spec:
        loc = [int(v) for v in spec["location"]]
        if dim == 2:
            # node grid coordinates, clamped to the far edge (problem.py:331-334)
            return int(node_id_2d(min(loc[0], nelx), min(loc[1], nely), nely))
        return int(node_id_3d(loc[0], loc[1], loc[2], nely, nelz))
    if "corner" in spec:
        name = spec["corner"]
        if dim == 2:
            if name not in _CORNERS_2D:
                raise ValueError(f"Unknown 2D corner '{name}'")
            return int(node_id_2d(*_CORNERS_2D[name](nelx, nely), nely))
        if name not in _CORNERS_3D:
            raise ValueError(f"Unknown 3D corner '{name}'")
        return int(node_id_3d(*_CORNERS_3D[name](nelx, nely, nelz), nely, nelz))
    raise ValueError("Point node spec requires 'location' or 'corner'")


def dof_offsets(labels, dim: int) -> list[int]:
    """DOF component offsets for labels like ['x', 'y'] (``problem.py:231-241``)."""
    comp = {"x": 0, "y": 1, "z": 2}
    out = []
    for lab in labels:
        if lab not in comp:
            raise ValueError(f"Unknown DOF label '{lab}'. Use 'x', 'y', or 'z'.")
        if comp[lab] >= dim:
            raise ValueError(f"DOF '{lab}' not valid for {dim}D problem")
        out.append(comp[lab])
    return out


def parse_fixed_dofs(bc_specs: dict, nelx, nely, nelz, dim) -> np.ndarray:
    """Sorted unique fixed DOF ids (``problem.py:176-198``)."""
    dofs: set[int] = set()
    for entry in bc_specs.get("fixed", []) or []:
        offs = dof_offsets(entry.get("dofs", ["x", "y"]), dim)
        for node in resolve_nodes(entry["nodes"], nelx, nely, nelz, dim):
            dofs.update(dim * node + o for o in offs)
    return np.array(sorted(dofs), dtype=int)


def parse_loads(load_specs, nelx, nely, nelz, dim) -> tuple[np.ndarray, np.ndarray]:
    """Per-node load vectors -> (dof ids, values); zero components dropped
    (``problem.py:201-228``). The force vector applies to every selected node."""
    dofs: list[int] = []
    vals: list[float] = []
    for entry in load_specs or []:
        vec = entry["force"]
        if len(vec) != dim:
            raise ValueError(f"force vector length {len(vec)} != dim {dim}")
        for node in resolve_nodes(entry["nodes"], nelx, nely, nelz, dim):
            for comp, val in enumerate(vec):
                if val != 0.0:
                    dofs.append(dim * node + comp)
                    vals.append(float(val))
    if not dofs:
        raise ValueError("No nonzero forces specified in problem definition")
    return np.array(dofs, dtype=int), np.array(vals, dtype=float)


def problem_from_dict(raw: dict) -> Problem:
    """Build a Problem from an already-parsed YAML mapping (``problem.py:87-164``)."""
    sec = {k: (raw.get(k) or {}) for k in
           ("problem", "domain", "material", "optimization", "solver", "output")}
    meta, dom, mat, opt, sol, out = (sec[k] for k in
                                     ("problem", "domain", "material", "optimization", "solver", "output"))

    dim = int(meta.get("dimension", 2))
    if dim not in (2, 3):
        raise ValueError(f"dimension must be 2 or 3, got {dim}")
    nelx, nely = int(dom["nelx"]), int(dom["nely"])
    nelz = int(dom.get("nelz", 1)) if dim == 3 else 1

    filter_type = opt.get("filter_type", "density")
    if filter_type not in ("density", "sensitivity"):
        raise ValueError(f"filter_type must be 'density' or 'sensitivity', got {filter_type}")
    precond = sol.get("preconditioner", "jacobi")
    if precond not in PRECONDITIONERS:
        raise ValueError(
            "preconditioner must be 'jacobi', 'block_jacobi', 'none', "
            f"'block_jacobi_2x2' or 'multigrid', got {precond}")
    precision = sol.get("precision", "fp64")
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision}")
    out_format = out.get("format", "png")
    if out_format not in ("png", "vtk"):
        raise ValueError(f"output format must be 'png' or 'vtk', got {out_format}")

    bcs = raw.get("boundary_conditions", {}) or {}
    fixed = parse_fixed_dofs(bcs, nelx, nely, nelz, dim)
    fdofs, fvals = parse_loads(raw.get("loads", []), nelx, nely, nelz, dim)

    return Problem(
        name=meta.get("name", "Topology Optimization"), dim=dim,
        nelx=nelx, nely=nely, nelz=nelz,
        E=float(mat.get("E", 1.0)), nu=float(mat.get("nu", 0.3)),
        E_min=float(mat.get("E_min", 1e-9)),
        fixed_dofs=fixed, force_dofs=fdofs, force_values=fvals,
        volfrac=float(opt.get("volume_fraction", 0.5)),
        penalty=float(opt.get("penalty", 3.0)),
        rmin=float(opt.get("filter_radius", 1.5)),
        filter_type=filter_type,
        max_iter=int(opt.get("max_iterations", 200)),
        tol=float(opt.get("tolerance", 0.01)),
        move_limit=float(opt.get("move_limit", 0.2)),
        max_cg_iter=int(sol.get("max_cg_iterations", 2000)),
        cg_tol=float(sol.get("cg_tolerance", 1e-6)),
        preconditioner=precond,
        output_dir=Path(out.get("directory", "results")),
        save_interval=int(out.get("save_interval", 10)),
        output_format=out_format,
        precision=precision,
    )


def load_problem(yaml_path: str | Path) -> Problem:
    """Parse and validate a YAML problem file (``problem.py:81-164``)."""
    with Path(yaml_path).open() as fh:
        raw = yaml.safe_load(fh)
    return problem_from_dict(raw or {})


def make_problem(nelx: int, nely: int, *, fixed: list, loads: list, name: str = "request",
                 nelz: int | None = None, **opts) -> Problem:
    """Programmatic constructor used for serving requests and benchmarks.

    ``fixed`` / ``loads`` use the YAML list formats of
    ``boundary_conditions.fixed`` and ``loads``; ``opts`` may hold any
    optimization/solver/material key under its YAML name
    (e.g. ``volume_fraction=0.3, filter_radius=6.0, precision='mixed'``).
    ``nelz`` makes a 3-D problem (``dimension: 3``).
    """
    sections = {"material": ("E", "nu", "E_min"),
                "optimization": ("volume_fraction", "penalty", "filter_radius", "filter_type",
                                 "max_iterations", "tolerance", "move_limit"),
                "solver": ("max_cg_iterations", "cg_tolerance", "preconditioner", "precision"),
                "output": ("directory", "save_interval", "format")}
    domain = {"nelx": int(nelx), "nely": int(nely)}
    if nelz is not None:
        domain["nelz"] = int(nelz)
    raw: dict = {"problem": {"name": name, "dimension": 2 if nelz is None else 3},
                 "domain": domain,
                 "boundary_conditions": {"fixed": fixed}, "loads": loads}
    for sec, keys in sections.items():
        raw[sec] = {k: opts.pop(k) for k in keys if k in opts}
    if opts:
        raise TypeError(f"unknown problem options: {sorted(opts)}")
    return problem_from_dict(raw)
