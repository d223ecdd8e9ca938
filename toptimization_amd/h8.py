"""3-D H8 path on libtopt (FP64): grid descriptor, op-level calls and the
SIMP loop for ``dim = 3``.

Replaces the reference's 3-D Taichi kernels (``fem/kernels.py:104-129,
179-212, 275-308``, ``fem/fused_kernels.py:59-94``,
``fem/preconditioner.py:33-44, 111-156``, ``optimizer/filter.py:75-97``)
and drives ``optimizer/simp.py:141-219`` for 3-D problems. Layout: node
``(ix*(nely+1) + iy)*(nelz+1) + iz``, three DOFs per node interleaved;
element ``(ex*nely + ey)*nelz + ez``; fixed DOFs as a bitmask. The PCG runs
on the device (``topt3_solve``); the OC update is the 2-D one over the
element array (``topt_oc_update`` with ``nely*nelz`` elements per column).
"""

from __future__ import annotations

import ctypes as C
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib, ops
from ._lib import TGrid3, check
from .engine import fixed_bitmask, force_vector, require_gpu, stream_ptr
from .material import compute_Ke_3d

PRECONDITIONERS_3D = {"none": _lib.PC_NONE, "jacobi": _lib.PC_JACOBI, "block_jacobi": _lib.PC_ELEMBLOCK}


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _s():
    return C.c_void_p(stream_ptr())


class Grid3:
    """One 3-D grid and material on the device: Ke (and the regularised
    inverse of the element-block preconditioner, ``fields.py:139-143``) in
    HBM, the ``topt_grid3`` descriptor and a call workspace."""

    def __init__(self, nelx: int, nely: int, nelz: int, Ke: np.ndarray | None = None, E_min: float = 1e-9,
                 penal: float = 3.0, device=None):
        self.device = device or require_gpu()
        self.nelx, self.nely, self.nelz = int(nelx), int(nely), int(nelz)
        Ke = compute_Ke_3d(1.0, 0.3) if Ke is None else np.asarray(Ke, dtype=np.float64)
        if Ke.shape != (24, 24):
            raise ValueError("H8 Ke must be 24x24")
        self.Ke_np = Ke
        self.alpha_reg = 0.1 * float(np.mean(np.diag(Ke)))
        self.Ke = torch.as_tensor(Ke.ravel().copy(), device=self.device)
        self.Ki = torch.as_tensor(np.linalg.inv(Ke + self.alpha_reg * np.eye(24)).ravel().copy(), device=self.device)
        self.desc = TGrid3(self.nelx, self.nely, self.nelz, 0, float(E_min), float(penal), self.Ke.data_ptr())
        nb = int(_lib.load().topt3_workspace_bytes(self.nelx, self.nely, self.nelz))
        if nb <= 0:
            raise ValueError("invalid 3-D grid")
        self.ws = torch.empty(nb, dtype=torch.uint8, device=self.device)

    @property
    def n_nodes(self) -> int:
        return (self.nelx + 1) * (self.nely + 1) * (self.nelz + 1)

    @property
    def n_dofs(self) -> int:
        return 3 * self.n_nodes

    @property
    def n_elem(self) -> int:
        return self.nelx * self.nely * self.nelz


def _f64(t: torch.Tensor) -> torch.Tensor:
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError("the 3-D path takes contiguous FP64 CUDA tensors")
    return t


def scale(g: Grid3, rho: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """s_e = E_min + rho_e^p (1 - E_min)."""
    s = torch.empty_like(rho) if out is None else out
    check(_lib.load().topt3_scale(C.byref(g.desc), _p(_f64(rho)), _p(_f64(s)), _s()), "topt3_scale")
    return s


def diagonal(g: Grid3, s: torch.Tensor, fix: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """diag K(s), fixed -> 1, clamp 1e-20 (``fem/kernels.py:104-129``)."""
    d = torch.empty(g.n_dofs, dtype=torch.float64, device=s.device) if out is None else out
    check(_lib.load().topt3_diagonal(C.byref(g.desc), _p(_f64(s)), _p(fix), _p(_f64(d)), _s()), "topt3_diagonal")
    return d


def matvec(g: Grid3, p: torch.Tensor, s: torch.Tensor, fix: torch.Tensor, Ap: torch.Tensor | None = None,
           with_dot: bool = False):
    """Ap = K(s) p with identity rows at fixed DOFs (``fem/kernels.py:179-212``);
    with_dot: also p.Ap (``fem/fused_kernels.py:59-94``)."""
    Ap = torch.empty_like(p) if Ap is None else Ap
    dot = torch.zeros(1, dtype=torch.float64, device=p.device) if with_dot else None
    check(_lib.load().topt3_matvec(C.byref(g.desc), _p(_f64(p)), _p(_f64(Ap)), _p(_f64(s)), _p(fix), _p(dot),
                                   _p(g.ws), _s()), "topt3_matvec")
    return (Ap, dot[0]) if with_dot else Ap


def elemblock_apply(g: Grid3, r: torch.Tensor, s: torch.Tensor, fix: torch.Tensor, z: torch.Tensor | None = None):
    """Element-block Jacobi z = M r; returns (z, r.z, r.r) (``fem/preconditioner.py:111-156``)."""
    z = torch.empty_like(r) if z is None else z
    out = torch.zeros(2, dtype=torch.float64, device=r.device)
    check(_lib.load().topt3_elemblock_apply(C.byref(g.desc), _p(g.Ki), g.alpha_reg, _p(_f64(r)), _p(_f64(z)),
                                            _p(_f64(s)), _p(fix), _p(out), _p(g.ws), _s()), "topt3_elemblock_apply")
    return z, out[0], out[1]


def sensitivity(g: Grid3, u: torch.Tensor, rho: torch.Tensor, dc: torch.Tensor | None = None):
    """(compliance 0-d tensor, dc) (``fem/kernels.py:275-308``)."""
    dc = torch.empty_like(rho) if dc is None else dc
    c = torch.zeros(1, dtype=torch.float64, device=u.device)
    check(_lib.load().topt3_sensitivity(C.byref(g.desc), _p(_f64(u)), _p(_f64(rho)), _p(_f64(dc)), None, _p(c),
                                        _p(g.ws), _s()), "topt3_sensitivity")
    return c[0], dc


def filter_apply(g: Grid3, rmin: float, mode: int, rho: torch.Tensor | None = None, dc: torch.Tensor | None = None,
                 out: torch.Tensor | None = None, Hs: torch.Tensor | None = None) -> torch.Tensor:
    """3-D density / dc-average / sensitivity filter (``optimizer/filter.py:106-171``
    on the tables of ``filter.py:75-97``)."""
    ref = rho if rho is not None else dc
    out = torch.empty(g.n_elem, dtype=torch.float64, device=ref.device) if out is None else out
    Hs = torch.empty_like(out) if Hs is None else Hs
    check(_lib.load().topt3_filter(g.nelx, g.nely, g.nelz, float(rmin), int(mode),
                                   _p(None if rho is None else _f64(rho)), _p(None if dc is None else _f64(dc)),
                                   _p(out), _p(Hs), _p(g.ws), _s()), "topt3_filter")
    return out


def filter_tables(g: Grid3, rmin: float):
    """The reference's padded 3-D (neighbors, weights, n_nb) tables."""
    width = (2 * math.ceil(rmin) + 1) ** 3
    n = g.n_elem
    nb = torch.empty(n * width, dtype=torch.int32, device=g.device)
    wt = torch.empty(n * width, dtype=torch.float32, device=g.device)
    cnt = torch.empty(n, dtype=torch.int32, device=g.device)
    check(_lib.load().topt3_filter_tables(g.nelx, g.nely, g.nelz, float(rmin), width, _p(nb), _p(wt), _p(cnt),
                                          _p(g.ws), _s()), "topt3_filter_tables")
    return nb.view(n, width), wt.view(n, width), cnt


def solve(g: Grid3, preconditioner: str, f: torch.Tensor, fix: torch.Tensor, s: torch.Tensor, u: torch.Tensor,
          warm: bool, tol: float, max_iter: int) -> tuple[int, int]:
    """PCG K(s) u = f on the device (``fem/solver.py:27-167``); returns
    (iterations in the reference's count, status)."""
    if preconditioner not in PRECONDITIONERS_3D:
        raise ValueError(f"3-D preconditioners are {sorted(PRECONDITIONERS_3D)}, got {preconditioner!r}")
    it = C.c_int32()
    st = C.c_int32()
    check(_lib.load().topt3_solve(C.byref(g.desc), PRECONDITIONERS_3D[preconditioner], _p(g.Ki), g.alpha_reg,
                                  _p(_f64(f)), _p(fix), _p(_f64(s)), _p(_f64(u)), int(bool(warm)), float(tol),
                                  int(max_iter), C.byref(it), C.byref(st), _p(g.ws), _s()), "topt3_solve")
    return int(it.value), int(st.value)


class Design3:
    """One 3-D design problem resident in HBM: rho, filtered rho, u, f, the
    fixed-DOF bitmask, sensitivities and filter row sums."""

    def __init__(self, problem, device=None):
        if problem.dim != 3:
            raise ValueError("Design3 takes a 3-D problem")
        if problem.preconditioner not in PRECONDITIONERS_3D:
            raise ValueError(f"3-D problems take the preconditioners {sorted(PRECONDITIONERS_3D)}")
        self.problem = problem
        self.g = Grid3(problem.nelx, problem.nely, problem.nelz, compute_Ke_3d(problem.E, problem.nu),
                       problem.E_min, problem.penalty, device)
        dev = self.g.device
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=dev)  # noqa: E731
        ne, nd = self.g.n_elem, self.g.n_dofs
        self.rho = torch.full((ne,), float(problem.volfrac), dtype=torch.float64, device=dev)
        self.rhof, self.dc, self.dcf, self.Hs, self.s = z(ne), z(ne), z(ne), z(ne), z(ne)
        self.u = z(nd)
        self.f = torch.as_tensor(force_vector(problem), device=dev)
        self.fix = torch.as_tensor(fixed_bitmask(nd, problem.fixed_dofs).view(np.int32), device=dev)
        self.iter = 0

    def iteration(self) -> dict:
        """One SIMP iteration (``optimizer/simp.py:144-197``)."""
        p, g = self.problem, self.g
        dens = p.filter_type == "density"
        if dens:
            filter_apply(g, p.rmin, _lib.FILTER_DENSITY, rho=self.rho, out=self.rhof, Hs=self.Hs)
        phys = self.rhof if dens else self.rho
        scale(g, phys, out=self.s)
        n_cg, status = solve(g, p.preconditioner, self.f, self.fix, self.s, self.u, warm=self.iter > 0,
                             tol=p.cg_tol, max_iter=p.max_cg_iter)
        c, _ = sensitivity(g, self.u, phys, dc=self.dc)
        if dens:
            filter_apply(g, p.rmin, _lib.FILTER_DCAVG, dc=self.dc, out=self.dcf, Hs=self.Hs)
        else:
            filter_apply(g, p.rmin, _lib.FILTER_SENS, rho=self.rho, dc=self.dc, out=self.dcf, Hs=self.Hs)
        change, vol, lam, steps = ops.oc_update(g.nelx, g.nely * g.nelz, self.rho, self.dcf, p.volfrac, p.move_limit)
        self.iter += 1
        return {"c": float(c.item()), "vol": vol, "change": change, "n_cg": n_cg, "cg_status": status,
                "lam": lam, "oc_steps": steps}


def run_3d(problem, *, verbose: bool = True, history: dict | None = None) -> np.ndarray:
    """The SIMP loop of ``optimizer/simp.py:141-231`` for a 3-D problem:
    same printout, stopping rule (``it > 5 and change < tol``) and
    ``density_final.npy`` (flat float32, e = (ex*nely + ey)*nelz + ez);
    returns rho as (nelx, nely, nelz) (``simp.py:309-312``)."""
    log = print if verbose else (lambda *a, **k: None)
    log(f"[toptimization] Backend: hip (gfx950), 3-D H8 path")
    log(f"[toptimization] Building mesh ({problem.nelx}x{problem.nely}x{problem.nelz})...")
    d = Design3(problem)
    out_dir = Path(problem.output_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    hist = history if history is not None else {}
    for k in ("c", "vol", "change", "n_cg", "oc_steps", "time"):
        hist.setdefault(k, [])
    log(f"\n[toptimization] Starting: {problem.name}")
    log(f"{'Iter':>5} {'Compliance':>14} {'Volume':>8} {'Change':>9} {'CG':>6} {'s/iter':>7}")
    log("-" * 55)
    for it in range(1, problem.max_iter + 1):
        t0 = time.perf_counter()
        st = d.iteration()
        el = time.perf_counter() - t0
        for k in ("c", "vol", "change", "n_cg", "oc_steps"):
            hist[k].append(st[k])
        hist["time"].append(el)
        log(f"{it:>5d} {st['c']:>14.6f} {st['vol']:>8.4f} {st['change']:>9.4f} {st['n_cg']:>6d} {el:>7.2f}")
        if it > 5 and st["change"] < problem.tol:
            log(f"\n[toptimization] Converged at iter {it} (change={st['change']:.6f})")
            break
    log("\n[toptimization] Done.")
    rho = d.rho.cpu().numpy()
    hist["u"] = d.u.cpu().numpy()
    np.save(out_dir / "density_final.npy", rho.astype(np.float32))
    return rho.reshape(problem.nelx, problem.nely, problem.nelz)
