"""Device-resident design problems and batched plans over libtopt.

``DesignState`` owns one problem's GPU workspace (a single torch uint8
tensor carved by ``topt_problem_layout``) and exposes torch views of the
fields the reference keeps as Taichi globals (``fem/fields.py:14-50``).
``Plan`` advances one or many states together: every kernel launch covers
all problems of the plan (ragged block table), and per-problem reductions
use only that problem's blocks, so a problem's results do not depend on
which other problems share the launch.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import BUF, TIterStats, TProblem, check
from .material import compute_Ke_2d
from .problem import Problem

PRECISION_IDS = {"fp64": _lib.F64, "mixed": _lib.F32}


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("toptimization_amd needs a ROCm GPU (MI355X); none is visible")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def force_vector(problem: Problem) -> np.ndarray:
    """FP64 load vector: duplicates summed, fixed entries zeroed (``fem/fields.py:151-155``)."""
    f = np.zeros(problem.n_dofs)
    np.add.at(f, np.asarray(problem.force_dofs, dtype=np.int64), np.asarray(problem.force_values, float))
    f[np.asarray(problem.fixed_dofs, dtype=np.int64)] = 0.0
    return f


def fixed_bitmask(n_dofs: int, fixed_dofs) -> np.ndarray:
    """Fixed-DOF bitmask, bit (d % 32) of word d // 32 (uint32 words)."""
    words = np.zeros((n_dofs + 31) // 32, dtype=np.uint32)
    d = np.asarray(fixed_dofs, dtype=np.int64)
    np.bitwise_or.at(words, d // 32, (np.uint32(1) << (d % 32).astype(np.uint32)))
    return words


@dataclass
class IterStats:
    compliance: float
    volume: float
    change: float
    lam: float
    n_cg: int
    cg_status: int
    oc_steps: int
    refines: int
    iter: int = 0
    done: int = 0


class DesignState:
    """One 2-D design problem resident in HBM.

    Fields (torch views, FP64): ``u`` (2*nodes), ``f``, ``rho``, ``rhof``,
    ``dc``, ``dcf``, ``Hs`` (elements); ``fix`` is the DOF bitmask.
    """

    def __init__(self, problem: Problem, precision: str | None = None,
                 preconditioner: str | None = None, delta: float | None = None,
                 device: torch.device | None = None, rho0: np.ndarray | None = None,
                 Ke: np.ndarray | None = None, f: np.ndarray | None = None):
        if problem.dim != 2:
            raise NotImplementedError("the GPU path covers 2-D Q4 problems only")
        self.problem = problem
        self.precision = precision or getattr(problem, "precision", "fp64")
        self.dtype_id = PRECISION_IDS[self.precision]
        pc_name = preconditioner or problem.preconditioner
        self.precond_id = _lib.PRECONDITIONER_IDS[pc_name]
        self.device = device or require_gpu()
        lib = _lib.load()
        nelx, nely = problem.nelx, problem.nely
        self.n_nodes = (nelx + 1) * (nely + 1)
        self.n_elem = nelx * nely
        offs = (C.c_int64 * len(BUF))()
        total = C.c_int64()
        check(lib.topt_problem_layout(nelx, nely, self.dtype_id, self.precond_id, float(problem.rmin),
                                      offs, C.byref(total)), "topt_problem_layout")
        self.offsets = list(offs)
        self.nbytes = int(total.value)
        self.ws = torch.zeros(self.nbytes, dtype=torch.uint8, device=self.device)

        nd = 2 * self.n_nodes
        self.u = self._view("u", torch.float64, nd)
        self.f = self._view("f", torch.float64, nd)
        self.fix = self._view("fix", torch.int32, (nd + 31) // 32)
        self.rho = self._view("rho", torch.float64, self.n_elem)
        self.rhof = self._view("rhof", torch.float64, self.n_elem)
        self.dc = self._view("dc", torch.float64, self.n_elem)
        self.dcf = self._view("dcf", torch.float64, self.n_elem)
        self.Hs = self._view("Hs", torch.float64, self.n_elem)

        f_np = force_vector(problem) if f is None else np.asarray(f, dtype=np.float64).ravel()
        self.f.copy_(torch.from_numpy(f_np))
        self.fix.copy_(torch.from_numpy(fixed_bitmask(nd, problem.fixed_dofs).view(np.int32)))
        if rho0 is None:
            self.rho.fill_(float(problem.volfrac))
        else:
            self.rho.copy_(torch.as_tensor(np.asarray(rho0, dtype=np.float64).ravel()))
        self.Ke = compute_Ke_2d(problem.E, problem.nu) if Ke is None else np.asarray(Ke, dtype=np.float64)
        self.desc = TProblem()
        d = self.desc
        d.nelx, d.nely = nelx, nely
        d.precond = self.precond_id
        d.max_cg_iter = int(problem.max_cg_iter)
        d.E_min, d.penal = float(problem.E_min), float(problem.penalty)
        d.cg_tol = float(problem.cg_tol)
        d.volfrac, d.move, d.rmin = float(problem.volfrac), float(problem.move_limit), float(problem.rmin)
        d.fnorm2 = float(f_np @ f_np)
        # reliable-update ratio of the mixed path: the FP64 residual is
        # recomputed when the recursive one has fallen by delta. Multigrid
        # solves converge in tens of iterations, so a refresh every three
        # decades costs less than the drift it prevents.
        if delta is None:
            delta = 1e-3 if self.precond_id == _lib.PC_MG else 0.1
        d.delta = float(delta)
        d.Ke[:] = [float(v) for v in self.Ke.ravel()]
        d.ws = self.ws.data_ptr()
        d.oc_tol = float(problem.tol)
        d.oc_max_iter = int(problem.max_iter)
        check(lib.topt_problem_init(C.byref(d), self.dtype_id, C.c_void_p(stream_ptr())),
              "topt_problem_init")

    def _view(self, name: str, dtype: torch.dtype, count: int) -> torch.Tensor:
        off = self.offsets[BUF[name]]
        nb = count * torch.empty((), dtype=dtype).element_size()
        return self.ws[off:off + nb].view(dtype)

    def buffer(self, name: str, dtype: torch.dtype, count: int) -> torch.Tensor:
        """Raw view of any workspace buffer (``topt.h`` TOPT_BUF_* names)."""
        return self._view(name, dtype, count)

    @property
    def vec_dtype(self) -> torch.dtype:
        return torch.float32 if self.dtype_id == _lib.F32 else torch.float64


class Plan:
    """A batch of DesignStates advanced by shared launches (``topt_plan_*``)."""

    def __init__(self, states: list[DesignState], graph: bool = False):
        """graph=True replays every full 32-iteration chunk of PCG launches
        from a HIP graph (topt_plan_set_graph); results are identical."""
        if not states:
            raise ValueError("empty plan")
        dt = {s.dtype_id for s in states}
        if len(dt) != 1:
            raise ValueError("all problems of a plan must share one precision")
        self.states = list(states)
        self.dtype_id = dt.pop()
        self._descs = (TProblem * len(states))(*[s.desc for s in states])
        handle = C.c_void_p()
        check(_lib.load().topt_plan_create(self._descs, len(states), self.dtype_id, C.byref(handle)),
              "topt_plan_create")
        self.handle = handle
        if graph:
            check(_lib.load().topt_plan_set_graph(self.handle, 1), "topt_plan_set_graph")

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.load().topt_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass

    def __len__(self) -> int:
        return len(self.states)

    def _stream(self):
        return C.c_void_p(stream_ptr())

    def filter(self, mode: int) -> None:
        check(_lib.load().topt_plan_filter(self.handle, mode, self._stream()), "topt_plan_filter")

    def solve(self, warm: int = 2, use_filtered: bool = True) -> tuple[list[int], list[int]]:
        n = len(self)
        it = (C.c_int32 * n)()
        st = (C.c_int32 * n)()
        check(_lib.load().topt_plan_solve(self.handle, int(warm), int(use_filtered), it, st,
                                          self._stream()), "topt_plan_solve")
        return list(it), list(st)

    def sensitivity(self, use_filtered: bool = True) -> None:
        check(_lib.load().topt_plan_sensitivity(self.handle, int(use_filtered), self._stream()),
              "topt_plan_sensitivity")

    def oc(self) -> None:
        check(_lib.load().topt_plan_oc(self.handle, self._stream()), "topt_plan_oc")

    def iteration(self, sensitivity_filter: bool = False) -> list[IterStats]:
        out = (TIterStats * len(self))()
        check(_lib.load().topt_plan_iteration(self.handle, int(bool(sensitivity_filter)), out,
                                              self._stream()), "topt_plan_iteration")
        return [_stats(o) for o in out]

    def advance(self, phase: list[int], sensitivity_filter: bool = False, max_chunks: int = 1 << 30):
        """Asynchronous SIMP iterations (topt_plan_advance): each problem runs
        its own OC iterations and PCG solves until one becomes done, until
        max_chunks chunks of 16 CG iterations, or until all are done.
        phase[i]: 0 start an OC iteration, 1 solving, 2 OC step pending,
        3 done; updated in place. Returns (stats of each problem's latest
        completed OC iteration, OC iterations completed per problem, their
        PCG iterations)."""
        n = len(self)
        if len(phase) != n:
            raise ValueError("one phase per problem")
        ph = (C.c_int32 * n)(*phase)
        out = (TIterStats * n)()
        done = (C.c_int32 * n)()
        cg = (C.c_int64 * n)()
        check(_lib.load().topt_plan_advance(self.handle, int(bool(sensitivity_filter)), int(max_chunks), ph, out,
                                            done, cg, self._stream()), "topt_plan_advance")
        phase[:] = list(ph)
        return [_stats(o) for o in out], list(done), list(cg)

    def time_slot(self, which: int, reps: int) -> float:
        """Average GPU ms of PCG slot kernel A (0) or B (1) over `reps`
        back-to-back launches (benchmark only; leaves the solve finished)."""
        ms = C.c_float()
        check(_lib.load().topt_plan_time_slot(self.handle, int(which), int(reps), C.byref(ms),
                                              self._stream()), "topt_plan_time_slot")
        return ms.value / reps

    def set_timing(self, enable: bool = True) -> None:
        """Accumulate the device time of each iteration() phase (HIP events on
        the launch stream); enabling resets the sums (benchmark support)."""
        check(_lib.load().topt_plan_set_timing(self.handle, int(bool(enable))), "topt_plan_set_timing")

    def timing(self) -> dict:
        """Summed device milliseconds of the iteration() phases since
        set_timing(True): 'setup' (density filter + solve setup), 'pcg' (all
        chunks of CG iterations), 'post' (sensitivity, dc filter, OC,
        statistics), and the number of iterations timed."""
        ms = (C.c_double * 3)()
        calls = C.c_int64()
        check(_lib.load().topt_plan_timing(self.handle, ms, C.byref(calls)), "topt_plan_timing")
        return {"setup": ms[0], "pcg": ms[1], "post": ms[2], "iterations": calls.value}

    def stats(self) -> list[IterStats]:
        out = (TIterStats * len(self))()
        check(_lib.load().topt_plan_stats(self.handle, out, self._stream()), "topt_plan_stats")
        return [_stats(o) for o in out]


def _stats(o: TIterStats) -> IterStats:
    return IterStats(o.compliance, o.volume, o.change, o.lam, o.n_cg, o.cg_status, o.oc_steps,
                     o.refines, o.iter, o.done)
