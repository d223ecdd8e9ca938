"""Op-level wrappers over libtopt: one call per reference Taichi kernel.

Each function takes torch CUDA tensors (node vectors interleaved (x, y),
element vectors in ``e = ex*nely+ey`` order) and runs the corresponding
``topt_*`` C entry point on the current stream. Used by the drop-in
``fem.kernels`` / ``optimizer.filter`` / ``optimizer.oc`` modules and by the
parity tests.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import TGrid, check
from .engine import fixed_bitmask, stream_ptr

_DT = {torch.float32: _lib.F32, torch.float64: _lib.F64}


def _p(t: torch.Tensor | None):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("libtopt ops take CUDA tensors")
    if not t.is_contiguous():
        raise ValueError("libtopt ops take contiguous tensors")
    return C.c_void_p(t.data_ptr())


def _s():
    return C.c_void_p(stream_ptr())


def _dtype_id(t: torch.Tensor) -> int:
    if t.dtype not in _DT:
        raise TypeError(f"unsupported dtype {t.dtype}")
    return _DT[t.dtype]


def make_grid(nelx: int, nely: int, Ke: np.ndarray, E_min: float = 1e-9, penal: float = 3.0) -> TGrid:
    g = TGrid()
    g.nelx, g.nely = int(nelx), int(nely)
    g.E_min, g.penal = float(E_min), float(penal)
    Ke = np.asarray(Ke, dtype=np.float64)
    if Ke.shape != (8, 8):
        raise ValueError("Q4 Ke must be 8x8")
    g.Ke[:] = [float(v) for v in Ke.ravel()]
    return g


def workspace(nelx: int, nely: int, device=None) -> torch.Tensor:
    nb = int(_lib.load().topt_op_workspace_bytes(int(nelx), int(nely)))
    if nb <= 0:
        raise ValueError("invalid grid")
    return torch.empty(nb, dtype=torch.uint8, device=device or "cuda")


def fix_mask(n_dofs: int, fixed_dofs, device=None) -> torch.Tensor:
    """Fixed-DOF bitmask tensor (int32 words) from a list of fixed DOF ids."""
    return torch.from_numpy(fixed_bitmask(n_dofs, fixed_dofs).view(np.int32)).to(device or "cuda")


def scale(g: TGrid, rho: torch.Tensor, dtype=torch.float64, out: torch.Tensor | None = None) -> torch.Tensor:
    """s_e = E_min + rho_e^p (1 - E_min) in dtype, into `out` when given."""
    if out is not None and (out.dtype != dtype or out.numel() != rho.numel()):
        raise ValueError("out must have rho's length and the requested dtype")
    s = torch.empty(rho.numel(), dtype=dtype, device=rho.device) if out is None else out
    check(_lib.load().topt_scale(C.byref(g), _DT[dtype], _p(rho), _p(s), _s()), "topt_scale")
    return s


def matvec(g: TGrid, p: torch.Tensor, s: torch.Tensor, fix: torch.Tensor, Ap: torch.Tensor | None = None,
           with_dot: bool = False):
    """Ap = K(s) p (fixed rows identity); with_dot -> also p.Ap as a 0-d FP64 tensor."""
    if p.dtype != s.dtype:
        raise TypeError("p and s must share a dtype")
    Ap = torch.empty_like(p) if Ap is None else Ap
    dot = torch.zeros(1, dtype=torch.float64, device=p.device) if with_dot else None
    ws = workspace(g.nelx, g.nely, p.device) if with_dot else None
    check(_lib.load().topt_matvec(C.byref(g), _dtype_id(p), _p(p), _p(Ap), _p(s), _p(fix), _p(dot), _p(ws),
                                  _s()), "topt_matvec")
    return (Ap, dot[0]) if with_dot else Ap


def diagonal(g: TGrid, s: torch.Tensor, fix: torch.Tensor) -> torch.Tensor:
    n = 2 * (g.nelx + 1) * (g.nely + 1)
    d = torch.empty(n, dtype=s.dtype, device=s.device)
    check(_lib.load().topt_diagonal(C.byref(g), _dtype_id(s), _p(s), _p(fix), _p(d), _s()), "topt_diagonal")
    return d


def block2_inverse(g: TGrid, s: torch.Tensor, fix: torch.Tensor):
    nn = (g.nelx + 1) * (g.nely + 1)
    dg = torch.empty(2 * nn, dtype=s.dtype, device=s.device)
    off = torch.empty(nn, dtype=s.dtype, device=s.device)
    check(_lib.load().topt_block2_inverse(C.byref(g), _dtype_id(s), _p(s), _p(fix), _p(dg), _p(off), _s()),
          "topt_block2_inverse")
    return dg, off


def elemblock_apply(g: TGrid, r: torch.Tensor, s: torch.Tensor, fix: torch.Tensor):
    z = torch.empty_like(r)
    out = torch.zeros(2, dtype=torch.float64, device=r.device)
    ws = workspace(g.nelx, g.nely, r.device)
    check(_lib.load().topt_elemblock_apply(C.byref(g), _dtype_id(r), _p(r), _p(z), _p(s), _p(fix), _p(out),
                                           _p(ws), _s()), "topt_elemblock_apply")
    return z, out[0], out[1]


def sensitivity(g: TGrid, u: torch.Tensor, rho: torch.Tensor):
    """(compliance 0-d tensor, dc) in FP64 (fem/kernels.py:234-272)."""
    dc = torch.empty_like(rho)
    c = torch.zeros(1, dtype=torch.float64, device=u.device)
    ws = workspace(g.nelx, g.nely, u.device)
    check(_lib.load().topt_sensitivity(C.byref(g), _p(u), _p(rho), _p(dc), None, _p(c), _p(ws), _s()),
          "topt_sensitivity")
    return c[0], dc


def filter_apply(nelx: int, nely: int, rmin: float, mode: int, rho: torch.Tensor | None = None,
                 dc: torch.Tensor | None = None) -> torch.Tensor:
    ref = rho if rho is not None else dc
    out = torch.empty(nelx * nely, dtype=torch.float64, device=ref.device)
    Hs = torch.empty_like(out)
    ws = workspace(nelx, nely, ref.device)
    check(_lib.load().topt_filter(int(nelx), int(nely), float(rmin), int(mode), _p(rho), _p(dc), _p(out),
                                  _p(Hs), _p(ws), _s()), "topt_filter")
    return out


def filter_tables(nelx: int, nely: int, rmin: float, device=None):
    """The reference's padded (neighbors, weights, n_nb) tables, expanded on the GPU."""
    import math
    width = (2 * math.ceil(rmin) + 1) ** 2
    dev = device or "cuda"
    n = nelx * nely
    nb = torch.empty(n * width, dtype=torch.int32, device=dev)
    wt = torch.empty(n * width, dtype=torch.float32, device=dev)
    cnt = torch.empty(n, dtype=torch.int32, device=dev)
    ws = workspace(nelx, nely, dev)
    check(_lib.load().topt_filter_tables(int(nelx), int(nely), float(rmin), width, _p(nb), _p(wt), _p(cnt),
                                         _p(ws), _s()), "topt_filter_tables")
    return nb.view(n, width), wt.view(n, width), cnt


def oc_update(nelx: int, nely: int, rho: torch.Tensor, dcf: torch.Tensor, volfrac: float, move: float):
    """In-place OC update of rho; returns (change, volume, lambda, steps) as floats."""
    out = torch.zeros(4, dtype=torch.float64, device=rho.device)
    ws = workspace(nelx, nely, rho.device)
    check(_lib.load().topt_oc_update(int(nelx), int(nely), _p(rho), _p(dcf), None, float(volfrac),
                                     float(move), _p(out), _p(ws), _s()), "topt_oc_update")
    o = out.cpu().tolist()
    return o[0], o[1], o[2], int(o[3])


def volume(nelx: int, nely: int, v: torch.Tensor) -> float:
    out = torch.zeros(1, dtype=torch.float64, device=v.device)
    ws = workspace(nelx, nely, v.device)
    check(_lib.load().topt_volume(int(nelx), int(nely), _p(v), _p(out), _p(ws), _s()), "topt_volume")
    return float(out.item())


def blas1(op: str, n: int, a: float = 0.0, x=None, y=None, z=None, w=None, v=None, d=None,
          fix: torch.Tensor | None = None, reduce: bool = False):
    """Vector op of ``topt_blas1`` (see include/topt.h); returns the 2-vector
    of reductions (device tensor) when reduce=True."""
    ref = next(t for t in (x, y, z, w, v) if t is not None)
    red = torch.zeros(2, dtype=torch.float64, device=ref.device) if reduce else None
    ws = (torch.empty(int(_lib.load().topt_blas1_workspace_bytes()), dtype=torch.uint8, device=ref.device)
          if reduce else None)
    check(_lib.load().topt_blas1(_lib.B1[op], _dtype_id(ref), int(n), float(a), _p(x), _p(y), _p(z), _p(w),
                                 _p(v), _p(d), _p(fix), _p(red), _p(ws), _s()), f"topt_blas1({op})")
    return red
