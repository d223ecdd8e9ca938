"""Continuous-batching design server on one GPU.

Requests (design problems of any size, boundary conditions and loads)
arrive over time. Every admitted problem keeps its rho / u / PCG / OC state
resident in its own device workspace. Each engine step advances all active
problems by one SIMP iteration with shared launches (one kernel per PCG slot
covers every active problem through the plan's ragged block table).
Finished problems leave, and waiting ones are admitted as slots and the HBM
budget allow, so new work overlaps work already on the GPU.

Per-problem results do not depend on the batch: every reduction of a
problem uses only that problem's blocks in a fixed order, so a design is
bitwise the same alone and in the mix (tested).

Stopping rule per design: ``it > 5 and change < tol`` or ``it == max_iter``
(reference ``optimizer/simp.py:217``).
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .engine import DesignState, Plan, PRECISION_IDS
from .problem import Problem, make_problem, problem_from_dict


@dataclass
class Request:
    rid: int
    problem: Problem
    arrival: float                     # seconds since the stream started


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    iterations: int
    compliance: float
    volume: float
    change: float
    n_cg: int
    rho: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def request_from_json(obj, default_id: int = 0) -> Request:
    """One serving request from its JSON form (a dict or a JSON string):

        {"id": 7, "arrival": 0.35,
         "problem": {<the YAML problem schema: problem, domain, material,
                     boundary_conditions, loads, optimization, solver>}}

    ``arrival`` (seconds from the stream start) defaults to 0 and ``id`` to
    ``default_id``; the problem mapping is validated exactly like a YAML file
    (reference ``problem.py:87-164``)."""
    import json
    if isinstance(obj, (str, bytes)):
        obj = json.loads(obj)
    if not isinstance(obj, dict) or "problem" not in obj:
        raise ValueError("a request is a JSON object with a 'problem' mapping")
    return Request(int(obj.get("id", default_id)), problem_from_dict(obj["problem"]),
                   float(obj.get("arrival", 0.0)))


def load_requests(path) -> list[Request]:
    """Requests from a JSON-lines file (one request object per line)."""
    out = []
    with open(path) as fh:
        for k, line in enumerate(fh):
            if line.strip():
                out.append(request_from_json(line, default_id=k))
    return out


def footprint_bytes(problem: Problem, precision: str, preconditioner: str | None = None) -> int:
    """Device bytes one resident problem needs (its workspace layout)."""
    offs = (C.c_int64 * len(_lib.BUF))()
    total = C.c_int64()
    pc = _lib.PRECONDITIONER_IDS[preconditioner or problem.preconditioner]
    _lib.check(_lib.load().topt_problem_layout(problem.nelx, problem.nely, PRECISION_IDS[precision], pc,
                                               float(problem.rmin), offs, C.byref(total)),
               "topt_problem_layout")
    return int(total.value)


@dataclass
class _Active:
    req: Request
    state: DesignState
    admitted: float
    it: int = 0
    n_cg: int = 0
    phase: int = 0          # asynchronous driver phase (topt_plan_advance)
    last: object = None     # statistics of the latest completed OC iteration


@dataclass
class ServerStats:
    results: list = field(default_factory=list)
    makespan: float = 0.0
    steps: int = 0
    max_active: int = 0

    def summary(self) -> dict:
        lat = np.array([r.latency for r in self.results]) if self.results else np.zeros(1)
        return {"designs": len(self.results), "makespan_s": self.makespan,
                "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)),
                "designs_per_hour": 3600.0 * len(self.results) / max(self.makespan, 1e-9),
                "engine_steps": self.steps, "max_active": self.max_active}


# preconditioners whose PCG iteration has three slots (A, B, C); a plan
# holds either these or the two-slot ones (engine.hip: topt_plan_create)
THREE_SLOT = ("block_jacobi", "multigrid")


def plan_class(problem: Problem, preconditioner: str | None = None) -> tuple[bool, bool]:
    """Key of the designs that can share one plan: (three-slot PCG,
    sensitivity filter). Designs of different classes are served by
    separate plans, advanced in turn."""
    pc = preconditioner or problem.preconditioner
    return (pc in THREE_SLOT, problem.filter_type == "sensitivity")


class DesignServer:
    """Admission by slot count and HBM budget.

    mode='async': every active design runs its own SIMP iterations
    (topt_plan_advance); a design whose PCG solve ends gets its OC step and
    its next solve at the next 16-iteration chunk boundary while the others
    keep solving, and a finished design leaves at once.
    mode='lockstep': one SIMP iteration of every active design per engine
    step (each step waits for the slowest solve of the batch); the steps
    replay HIP graphs: the filter and solve setup, the PCG chunks of a
    multigrid solve, and the sensitivity, dc filter, OC update and
    statistics.
    mode='auto' (default): lockstep for multigrid designs, whose solves are
    short and even (512-request stream: p50 1.6 s, p99 2.0 s against 2.9 s
    and 6.1 s async), async for the others, whose solves vary by thousands
    of iterations. All modes give bitwise the same designs.

    Requests may name any preconditioner and filter type. Designs share a
    plan within their class (``plan_class``); a stream that mixes classes
    runs one plan per class, advanced in turn, with the same results."""

    def __init__(self, max_active: int = 64, hbm_budget: int = 32 << 30, precision: str = "mixed",
                 preconditioner: str | None = None, keep_rho: bool = False, mode: str = "auto",
                 admit_chunks: int = 8):
        if mode not in ("auto", "async", "lockstep"):
            raise ValueError("mode must be 'auto', 'async' or 'lockstep'")
        if preconditioner is not None and preconditioner not in _lib.PRECONDITIONER_IDS:
            raise ValueError(f"unknown preconditioner {preconditioner!r}")
        self.max_active = int(max_active)
        self.hbm_budget = int(hbm_budget)
        self.precision = precision
        self.preconditioner = preconditioner
        self.keep_rho = keep_rho
        self.mode = mode
        self.admit_chunks = int(admit_chunks)   # chunks per call while requests wait

    def _lockstep(self, problem: Problem) -> bool:
        """Whether the plan of this design advances in lockstep."""
        if self.mode != "auto":
            return self.mode == "lockstep"
        return (self.preconditioner or problem.preconditioner) == "multigrid"

    def _admit(self, pending, groups, used, t, clock) -> tuple[int, set]:
        """Admit arrived requests in order while slots and memory allow;
        returns the new memory use and the classes whose membership grew."""
        touched = set()
        n_active = sum(len(g) for g in groups.values())
        while pending and n_active < self.max_active:
            r = pending[0]
            if clock == "wall" and r.arrival > t:
                break
            need = footprint_bytes(r.problem, self.precision, self.preconditioner)
            if n_active and used + need > self.hbm_budget:
                break
            pending.pop(0)
            st = DesignState(r.problem, precision=self.precision, preconditioner=self.preconditioner)
            key = (*plan_class(r.problem, self.preconditioner), self._lockstep(r.problem))
            groups.setdefault(key, []).append(_Active(r, st, t))
            touched.add(key)
            used += need
            n_active += 1
        return used, touched

    def run(self, requests: list[Request], clock: str = "wall", progress=None) -> ServerStats:
        """Serve the stream. clock='wall': arrivals are real time offsets from
        the start; clock='instant': every request is available at t=0.
        progress(msg) is called whenever a design finishes."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        groups: dict[tuple, list[_Active]] = {}
        plans: dict[tuple, Plan] = {}
        stats = ServerStats()
        used = 0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731

        def finish(a: _Active, s, t):
            pr = a.req.problem
            stats.results.append(Result(
                a.req.rid, a.req.arrival, a.admitted, t, a.it, s.compliance, s.volume, s.change,
                a.n_cg, a.state.rho.cpu().numpy() if self.keep_rho else None))
            if progress is not None:
                r = stats.results[-1]
                progress(f"design {r.rid} ({pr.nelx}x{pr.nely}) done after {r.iterations} it, "
                         f"latency {r.latency:.2f} s, {len(stats.results)} finished, "
                         f"{len(pending)} waiting")
            return footprint_bytes(pr, self.precision, self.preconditioner)

        try:
            while pending or any(groups.values()):
                used, touched = self._admit(pending, groups, used, now(), clock)
                for key in touched:
                    if key in plans:
                        plans.pop(key).close()
                live = [k for k, g in groups.items() if g]
                if not live:
                    if pending and clock == "wall":
                        time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                    continue
                n_active = sum(len(groups[k]) for k in live)
                stats.max_active = max(stats.max_active, n_active)
                # while requests wait for a free slot, or other classes wait
                # for their turn, return often
                short = (bool(pending) and n_active < self.max_active) or len(live) > 1
                for key in live:
                    members = groups[key]
                    if key not in plans:
                        plans[key] = Plan([a.state for a in members])
                    plan = plans[key]
                    if not key[2]:   # async
                        phase = [a.phase for a in members]
                        out, done_now, cg_now = plan.advance(
                            phase, sensitivity_filter=key[1],
                            max_chunks=self.admit_chunks if short else 256)
                        t = now()
                        keep = []
                        for a, ph, s, dn, cg in zip(members, phase, out, done_now, cg_now):
                            a.phase = ph
                            a.it += dn
                            a.n_cg += cg
                            if dn:
                                a.last = s
                            if ph == 3:
                                used -= finish(a, a.last, t)
                            else:
                                keep.append(a)
                    else:
                        out = plan.iteration(sensitivity_filter=key[1])
                        t = now()
                        keep = []
                        for a, s in zip(members, out):
                            a.it += 1
                            a.n_cg += s.n_cg
                            pr = a.req.problem
                            if a.it >= pr.max_iter or (a.it > 5 and s.change < pr.tol):
                                used -= finish(a, s, t)
                            else:
                                keep.append(a)
                    if len(keep) != len(members):
                        groups[key] = keep
                        plans.pop(key).close()
                stats.steps += 1
                if progress is not None and stats.steps % 25 == 1:
                    progress(f"step {stats.steps} at {now():.1f} s: "
                             f"{sum(len(g) for g in groups.values())} active, "
                             f"{len(stats.results)} finished, {len(pending)} waiting")
        finally:
            for p in plans.values():
                p.close()
        stats.makespan = now()
        stats.results.sort(key=lambda r: r.rid)
        return stats


def synthetic_stream(n: int = 512, rate: float = 20.0, seed: int = 0, nelx_range=(64, 768),
                     max_iter: int = 100, tol: float = 0.01, **solver) -> list[Request]:
    """Seeded request stream: Poisson arrivals at `rate`/s; nelx uniform over
    multiples of 16 in nelx_range; nely = round(nelx * aspect), aspect in
    {1/4, 1/3, 1/2} (round half to even); BC type uniform over cantilever /
    MBB / bridge; seeded load node and angle; volfrac U[0.3, 0.5];
    rmin = 0.008 nelx + 1."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    sizes = np.arange(nelx_range[0], nelx_range[1] + 1, 16)
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        nelx = int(rng.choice(sizes))
        aspect = float(rng.choice([0.25, 1.0 / 3.0, 0.5]))
        nely = max(1, int(round(nelx * aspect)))
        kind = ["cantilever", "mbb", "bridge"][int(rng.integers(0, 3))]
        th = float(rng.uniform(0.0, 2.0 * np.pi))
        fx, fy = float(np.cos(th)), float(np.sin(th))
        if kind == "cantilever":
            fixed = [{"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x", "y"]}]
            loads = [{"nodes": {"type": "point", "location": [nelx, int(rng.integers(0, nely + 1))]},
                      "force": [fx, fy]}]
        elif kind == "mbb":
            fixed = [{"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x"]},
                     {"nodes": {"type": "point", "corner": "bottom-right"}, "dofs": ["y"]}]
            loads = [{"nodes": {"type": "point", "location": [int(rng.integers(0, nelx + 1)), nely]},
                      "force": [0.0, -1.0]}]
        else:
            fixed = [{"nodes": {"type": "point", "corner": "bottom-left"}, "dofs": ["x", "y"]},
                     {"nodes": {"type": "point", "corner": "bottom-right"}, "dofs": ["x", "y"]}]
            loads = [{"nodes": {"type": "edge", "edge": "top"}, "force": [0.0, -1.0 / (nelx + 1)]}]
        prob = make_problem(nelx, nely, fixed=fixed, loads=loads, name=f"{kind}-{i}",
                            volume_fraction=float(rng.uniform(0.3, 0.5)),
                            filter_radius=0.008 * nelx + 1.0, max_iterations=max_iter, tolerance=tol,
                            **solver)
        out.append(Request(i, prob, t))
    return out


def batch_cantilevers(n: int = 64, nelx: int = 128, nely: int = 64, seed: int = 0, max_iter: int = 100,
                      **solver) -> list[Problem]:
    """Independent cantilevers for one ragged batch: left edge clamped, unit
    point load on a seeded right-edge node with angle U[0, 2 pi), volfrac
    U[0.3, 0.6], rmin 2.5, p 3 (problem i uses default_rng(seed + i))."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed + i)
        iy = int(rng.integers(0, nely + 1))
        th = float(rng.uniform(0.0, 2.0 * np.pi))
        vf = float(rng.uniform(0.3, 0.6))
        out.append(make_problem(
            nelx, nely,
            fixed=[{"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x", "y"]}],
            loads=[{"nodes": {"type": "point", "location": [nelx, iy]},
                    "force": [float(np.cos(th)), float(np.sin(th))]}],
            name=f"cantilever-{i}", volume_fraction=vf, filter_radius=2.5, penalty=3.0,
            max_iterations=max_iter, tolerance=0.0, **solver))
    return out

