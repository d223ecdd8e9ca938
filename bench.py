"""Benchmark: SIMP OC iterations/s on one MI355X (1200x400 bridge, mixed precision).

A step is one OC iteration of the SIMP loop (density filter -> PCG ->
sensitivity -> dc filter -> OC bisection) over the 1200x400 Q4 bridge:
bottom corners pinned, total unit load spread over the top row, volfrac 0.3,
rmin 6, p 3, PCG rtol 1e-6 with FP32 vectors, FP64 dots and FP64 residual
refresh, geometric multigrid V(1,1) preconditioner. Inputs live in HBM
before the timed region.

Also reported (same JSON line): PCG iterations/s, the device time of the
iteration's phases (HIP events), the roofline of the multigrid CG iteration
(11 launches: PCG slots A, B, C and the V-cycle; algorithmic bytes over
device time per CG iteration, HBM traffic from the committed rocprofv3
summary), slot A alone, the whole 200-iteration run, matvec GDOF/s, the
plain FP64 GPU path on the same iterations, the 2x2 block-Jacobi path, the
compliance agreement of the two precisions on one density, and the C
restatement of the reference algorithm (FP32, Jacobi, as the reference) on
the host cores as the CPU baseline. The line's "configs" object carries the
other workloads measured in the same run: the 300x100 MBB, the batch of 64
cantilevers and the 512-request serving stream (--quick leaves them out).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
       python bench.py --mbb        (300x100 MBB, 100 OC iterations, three paths)
       python bench.py --batch      (64 x 128x64 cantilevers, batched vs one at a time)
       python bench.py --serving    (512-request stream: designs/hour, p50/p99 latency)
       python bench.py --h8         (3-D 96x48x48 cantilever and the 3-D matvec)
"""

from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NELX, NELY = 1200, 400
# rocprofv3 summaries this bench takes its measured HBM traffic from: each
# holds a line "traffic_bytes_per_<unit>: N" (FETCH_SIZE and WRITE_SIZE of
# the unit's launches, scaled by the copy calibration of the same runs)
PROFILE_MG = os.path.join(ROOT, "profiles", "multigrid_bridge1200x400_mixed.md")
PROFILE_SLOTS = os.path.join(ROOT, "profiles", "pcg_slots_bridge1200x400_mixed.md")
PROFILE_MV = os.path.join(ROOT, "profiles", "matvec_modal_8192x4096.md")


def profile_traffic(path, unit):
    """Measured HBM bytes per `unit` from a committed rocprofv3 summary, or
    None when the summary or its line is absent."""
    import re
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(rf"traffic_bytes_per_{unit}:\s*([0-9]+)", text)
    return int(m.group(1)) if m else None


def mg_levels(nelx, nely, dense=128):
    """Multigrid hierarchy: 2x2 agglomeration with ceil halving down to a
    coarsest level of at most `dense` DOFs."""
    lv = [(nelx, nely)]
    while True:
        nx, ny = (lv[-1][0] + 1) // 2, (lv[-1][1] + 1) // 2
        lv.append((nx, ny))
        if 2 * (nx + 1) * (ny + 1) <= dense:
            return lv


def mg_cg_bytes(nelx, nely, ts=4, tail=1024, tail_sum=2560):
    """Algorithmic bytes of one multigrid PCG iteration of the mixed path
    (vectors in ts-byte words, two per node; the FP64 solution u; node
    records of 40 words (the two rows of the assembled level operator and
    1/diag) at levels 1 .. L-1; the coarsest inverse in ts-byte words):
      slot A        read z, p, u; write p, Ap, u; read s
      slot B + down 0  read r, Ap, 1/diag, s; write r (the other buffer), b_1
      down l, up l  (1 <= l < tail level) read b_l, 1/diag_l, records_l,
                    [x_{l+1}]; write b_{l+1} / x_l
      tail          read b of its first level, the records of its sweep
                    levels and the coarsest inverse; write x of its first
                    level
      up 0 + slot C read r, 1/diag, x_1, s; write z"""
    lv = mg_levels(nelx, nely)
    L = len(lv) - 1
    nn = [(x + 1) * (y + 1) for x, y in lv]
    ne = [x * y for x, y in lv]
    lt = 1
    while lt < L and (nn[lt] > tail or sum(nn[lt:]) > tail_sum):
        lt += 1
    v = 2 * ts
    rec = 40 * ts
    b = 4 * v * nn[0] + 32 * nn[0] + ts * ne[0]          # slot A
    b += 4 * v * nn[0] + ts * ne[0] + v * nn[1]             # slot B + level-0 down sweep + restriction
    for l in range(1, lt):
        b += 2 * v * nn[l] + rec * nn[l] + v * nn[l + 1]       # down sweep + restriction
        b += 3 * v * nn[l] + rec * nn[l] + v * nn[l + 1]       # up sweep
    b += v * nn[lt] + sum(rec * nn[l] for l in range(lt, L)) + ts * (2 * nn[L]) ** 2 \
        + v * nn[lt]                                        # tail
    b += 3 * v * nn[0] + ts * ne[0] + v * nn[1]             # level-0 up sweep + slot C
    return b


def bridge(nelx, nely, precision, preconditioner="block_jacobi_2x2", tol=1e-6):
    from toptimization_amd.problem import make_problem
    return make_problem(
        nelx, nely,
        fixed=[{"nodes": {"type": "point", "corner": "bottom-left"}, "dofs": ["x", "y"]},
               {"nodes": {"type": "point", "corner": "bottom-right"}, "dofs": ["x", "y"]}],
        loads=[{"nodes": {"type": "edge", "edge": "top"}, "force": [0.0, -1.0 / (nelx + 1)]}],
        name="bridge", volume_fraction=0.3, filter_radius=6.0, penalty=3.0, max_iterations=200,
        tolerance=0.0, cg_tolerance=tol, max_cg_iterations=50000, preconditioner=preconditioner,
        precision=precision)


MAIN_PC = "multigrid"   # preconditioner of the headline path


def run_path(precision, warmup, steps, log, preconditioner=MAIN_PC):
    import torch
    from toptimization_amd.engine import DesignState, Plan
    prob = bridge(NELX, NELY, precision, preconditioner=preconditioner)
    st = DesignState(prob)
    plan = Plan([st])
    tag = f"{precision} {preconditioner}"
    for k in range(warmup):
        s = plan.iteration()[0]
        log(f"  [{tag}] warmup it {k + 1}: c={s.compliance:.6e} n_cg={s.n_cg} refines={s.refines}")
    # Python's cyclic garbage collector is kept out of the timed loop, as
    # timeit does: a full collection over torch's objects stalls the host for
    # milliseconds while the GPU waits for the next launch
    gc.collect()
    gc.disable()
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [plan.iteration()[0] for _ in range(steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    finally:
        gc.enable()
    ncg = sum(s.n_cg for s in stats)
    log(f"  [{tag}] timed {steps} it: {dt:.3f} s ({steps / dt:.2f} OC it/s), {ncg} CG iterations, "
        f"last c={stats[-1].compliance:.6e}")
    # the same OC iterations again, from a fresh state, with the phases timed
    # by HIP events on the launch stream (event reads between iterations, so
    # this pass is not the headline); the CG counts must repeat exactly
    st_t = DesignState(prob)
    plan_t = Plan([st_t])
    for _ in range(warmup):
        plan_t.iteration()
    plan_t.set_timing(True)
    stats_t = [plan_t.iteration()[0] for _ in range(steps)]
    tm = plan_t.timing()
    plan_t.close()
    tm["n_cg"] = sum(s.n_cg for s in stats_t)
    if tm["n_cg"] != ncg:
        raise RuntimeError(f"timed replay took {tm['n_cg']} CG iterations against {ncg}")
    log(f"  [{tag}] device time per OC iteration: setup {tm['setup'] / steps:.3f} ms, PCG "
        f"{tm['pcg'] / steps:.3f} ms ({1e3 * tm['pcg'] / max(tm['n_cg'], 1):.1f} us per CG iteration), "
        f"post {tm['post'] / steps:.3f} ms")
    return {"seconds": dt, "n_cg": ncg, "stats": stats, "state": st, "plan": plan, "problem": prob,
            "timing": tm}


def full_run(log, iters=200):
    """Config 2 as stated: all 200 OC iterations of the bridge from rho = 0.3
    (mixed, multigrid), wall time of the whole loop."""
    import torch
    from toptimization_amd.engine import DesignState, Plan
    prob = bridge(NELX, NELY, "mixed", preconditioner=MAIN_PC)
    st = DesignState(prob)
    plan = Plan([st])
    gc.collect()
    gc.disable()   # as in run_path
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ncg = 0
        for k in range(iters):
            s = plan.iteration()[0]
            ncg += s.n_cg
            if k % 50 == 49:
                log(f"  [full run] {k + 1} OC iterations, c={s.compliance:.6e}")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    finally:
        gc.enable()
    plan.close()
    log(f"  [full run] {iters} OC iterations in {dt:.3f} s ({iters / dt:.1f} OC it/s), {ncg} CG iterations")
    return {"seconds": dt, "n_cg": ncg, "iterations": iters, "final_compliance": s.compliance}


def same_rho_compliance(rho, log, preconditioner=MAIN_PC):
    """Solve the same density with the mixed and the plain FP64 path; compare c."""
    from toptimization_amd.engine import DesignState, Plan
    out = {}
    for prec in ("mixed", "fp64"):
        st = DesignState(bridge(NELX, NELY, prec, preconditioner=preconditioner), rho0=rho)
        plan = Plan([st])
        plan.filter(0)
        it, status = plan.solve(warm=0, use_filtered=True)
        plan.sensitivity(use_filtered=True)
        out[prec] = (plan.stats()[0].compliance, it[0])
        plan.close()
    rel = abs(out["mixed"][0] - out["fp64"][0]) / abs(out["fp64"][0])
    log(f"  same-rho compliance: mixed {out['mixed'][0]:.12e} ({out['mixed'][1]} CG) vs fp64 "
        f"{out['fp64'][0]:.12e} ({out['fp64'][1]} CG): rel {rel:.2e}")
    return rel


def kernel_roofline(plan, nn, ne, log, reps=200):
    """Average duration of PCG slot A (k_slot_a<float>) over back-to-back
    launches, timed with HIP events on the launch stream."""
    ms = plan.time_slot(0, reps)
    # FP32 slot A per node: read z, p_old (8 B each), write p_new, Ap (8 B each),
    # read and write the FP64 solution u (deferred u += alpha p: 16 B each);
    # per element: read s_e (4 B)
    bytes_a = 64 * nn + 4 * ne
    gbs = bytes_a / (ms * 1e-3) / 1e9
    log(f"  slot A: {ms * 1e3:.2f} us/launch, {bytes_a / 1e6:.2f} MB -> {gbs:.0f} GB/s")
    return ms, bytes_a, gbs


def matvec_gdofs(nelx, nely, log, reps=200, dtype="f32"):
    """Op-level K(s) p (no dot): `reps` launches captured in one HIP graph on
    the launch stream, replayed three times, the fastest replay per launch
    (torch events around the replay); rho ~ U[0.001, 1] (seed 0),
    p ~ N(0, 1) (seed 1), node 0 fixed."""
    import numpy as np
    import torch
    from toptimization_amd import ops
    from toptimization_amd.material import compute_Ke_2d
    nn = (nelx + 1) * (nely + 1)
    tdt = torch.float32 if dtype == "f32" else torch.float64
    g = ops.make_grid(nelx, nely, compute_Ke_2d(1.0, 0.3), 1e-9, 3.0)
    rho = torch.as_tensor(np.random.default_rng(0).uniform(0.001, 1.0, nelx * nely), device="cuda")
    s = ops.scale(g, rho, tdt)
    p = torch.as_tensor(np.random.default_rng(1).normal(size=2 * nn), dtype=tdt, device="cuda")
    fix = ops.fix_mask(2 * nn, [0, 1])
    Ap = torch.empty_like(p)
    for _ in range(3):
        ops.matvec(g, p, s, fix, Ap=Ap)
    torch.cuda.synchronize()
    st = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(st):
        with torch.cuda.graph(graph, stream=st):
            for _ in range(reps):
                ops.matvec(g, p, s, fix, Ap=Ap)
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = float("inf")
    for _ in range(3):
        e0.record()
        graph.replay()
        e1.record()
        e1.synchronize()
        ms = min(ms, e0.elapsed_time(e1) / reps)
    del graph
    gdofs = 2 * nn / (ms * 1e-3) / 1e9
    ws = 4 if dtype == "f32" else 8
    bytes_mv = ws * (2 * 2 * nn + nelx * nely)
    log(f"  matvec {dtype} {nelx}x{nely}: {ms * 1e3:.2f} us, {gdofs:.2f} GDOF/s, "
        f"{bytes_mv / (ms * 1e-3) / 1e9:.0f} GB/s algorithmic")
    return ms, gdofs, bytes_mv / (ms * 1e-3) / 1e9


def cpu_baseline(log, budget_s=25.0):
    """C restatement of the reference (FP32, Jacobi, host cores) on the first
    OC iterations of the same workload, bounded to about budget_s."""
    from oracle import build as obuild
    from oracle import c_oracle
    obuild.build() if not c_oracle.LIB.exists() else None
    cores = len(os.sched_getaffinity(0))
    c_oracle.set_threads(cores)
    prob = bridge(NELX, NELY, "fp64", preconditioner="jacobi")
    n_it = 1
    res = c_oracle.simp(prob, "f32", max_iter=n_it, preconditioner="jacobi")
    if res["seconds"] < budget_s / 3:
        n_it = max(1, min(4, int(budget_s / max(res["seconds"], 1e-3))))
        res = c_oracle.simp(prob, "f32", max_iter=n_it, preconditioner="jacobi")
    rate = res["iterations"] / res["seconds"]
    log(f"  CPU (C restatement, FP32 Jacobi, {cores} threads): {res['iterations']} OC it in "
        f"{res['seconds']:.2f} s, CG {res['n_cg']}")
    return {"value": rate, "unit": "OC it/s", "cores": cores, "kind": "port",
            "sample": f"first {res['iterations']} OC iteration(s) of the {NELX}x{NELY} bridge from rho=0.3, "
                      f"FP32 Jacobi PCG rtol 1e-6 ({sum(res['n_cg'])} CG iterations), C+OpenMP "
                      f"restatement of the reference algorithm"}


def batch_bench(log, n=64, iters=100, seq_sample=None, preconditioner=MAIN_PC):
    """64 independent 128x64 cantilevers (seeded loads, volfrac U[0.3, 0.6],
    rmin 2.5, mixed FP32/FP64 PCG rtol 1e-6, multigrid by default) advanced
    together by one plan (one launch per kernel for all of them) against the
    same problems solved one at a time. Every problem's iterates are the same
    either way (per-problem reductions in a fixed order)."""
    import torch
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import batch_cantilevers
    probs = batch_cantilevers(n, max_iter=iters, cg_tolerance=1e-6, max_cg_iterations=50000,
                              preconditioner=preconditioner, precision="mixed")
    states = [DesignState(p) for p in probs]
    plan = Plan(states)
    plan.iteration()   # warm-up: first iteration (setup, cold caches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ncg = 0
    for k in range(iters - 1):
        ncg += sum(s.n_cg for s in plan.iteration())
        if k % 20 == 19:
            log(f"  batch: {k + 2} OC iterations done")
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    plan.close()
    batched = n * (iters - 1) / tb
    log(f"  batch {n} x 128x64: {iters - 1} OC it each in {tb:.2f} s -> {batched:.1f} OC it/s "
        f"({ncg} CG iterations)")
    m = n if seq_sample is None else min(n, seq_sample)
    ts = 0.0
    for i, p in enumerate(probs[:m]):
        if i % 8 == 0:
            log(f"  sequential: problem {i}")
        st = DesignState(p)
        pl = Plan([st])
        pl.iteration()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters - 1):
            pl.iteration()
        torch.cuda.synchronize()
        ts += time.perf_counter() - t0
        pl.close()
    seq = m * (iters - 1) / ts
    log(f"  sequential ({m} problems): {seq:.1f} OC it/s")
    return {"metric": f"batched OC it/s, {n} x 128x64 cantilevers", "value": round(batched, 2),
            "preconditioner": preconditioner,
            "unit": "OC it/s", "sequential_oc_its_per_s": round(seq, 2),
            "sequential_sample": m, "speedup": round(batched / seq, 2), "oc_iterations_each": iters - 1,
            "cg_iterations_batched": ncg, "dtype": "f32 vectors + f64 dots/refresh", "data": "synthetic"}


def mbb_bench(log, iters=100):
    """300x100 half MBB beam (left edge x-roller, bottom-right y-roller, unit
    load at the top-left corner), volfrac 0.5, rmin 4, p 3, 100 OC iterations
    from rho = 0.5: the mixed FP32/FP64 path with 2x2 block Jacobi (rtol 1e-6)
    against the plain FP64 path (same preconditioner), against mixed Jacobi
    (CG count of the two preconditioners) and against mixed multigrid (the
    drop-in preconditioner of the bridge line)."""
    import numpy as np
    import torch
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.problem import make_problem

    def mbb(prec, pc):
        return make_problem(
            300, 100,
            fixed=[{"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x"]},
                   {"nodes": {"type": "point", "corner": "bottom-right"}, "dofs": ["y"]}],
            loads=[{"nodes": {"type": "point", "corner": "top-left"}, "force": [0.0, -1.0]}],
            name="mbb", volume_fraction=0.5, filter_radius=4.0, max_iterations=iters, tolerance=0.0,
            cg_tolerance=1e-6, max_cg_iterations=50000, preconditioner=pc, precision=prec)

    out = {}
    for key, prec, pc in (("mixed", "mixed", "block_jacobi_2x2"), ("fp64", "fp64", "block_jacobi_2x2"),
                          ("mixed_jacobi", "mixed", "jacobi"), ("mixed_multigrid", "mixed", "multigrid")):
        st = DesignState(mbb(prec, pc))
        plan = Plan([st])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [plan.iteration()[0] for _ in range(iters)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        plan.close()
        ncg = sum(x.n_cg for x in stats)
        out[key] = {"seconds": dt, "n_cg": ncg, "c": np.array([x.compliance for x in stats])}
        log(f"  [mbb 300x100 {key}] {iters} it in {dt:.3f} s: {iters / dt:.1f} OC it/s, {ncg} CG "
            f"({ncg / dt:.0f} CG it/s), final c={stats[-1].compliance:.9e}")
    rel = np.abs(out["mixed"]["c"] - out["fp64"]["c"]) / np.abs(out["fp64"]["c"])
    nd, ne = 2 * 301 * 101, 300 * 100
    # PCG-iteration algorithmic bytes of the mixed 2x2 block-Jacobi path:
    # slot A 64 N_n + 4 N_e, slot B 44 N_n (N_n = N_d / 2 nodes)
    b_it = 64 * nd // 2 + 4 * ne + 44 * nd // 2
    m = out["mixed"]
    return {"metric": "OC it/s (300x100 MBB, 100 iterations)", "value": round(iters / m["seconds"], 2),
            "unit": "OC it/s", "pcg_its_per_s": round(m["n_cg"] / m["seconds"], 1), "cg_total": m["n_cg"],
            "fp64_oc_its_per_s": round(iters / out["fp64"]["seconds"], 2), "fp64_cg_total": out["fp64"]["n_cg"],
            "mixed_over_fp64": round(out["fp64"]["seconds"] / m["seconds"], 3),
            "jacobi_cg_total": out["mixed_jacobi"]["n_cg"],
            "jacobi_oc_its_per_s": round(iters / out["mixed_jacobi"]["seconds"], 2),
            "multigrid_oc_its_per_s": round(iters / out["mixed_multigrid"]["seconds"], 2),
            "multigrid_cg_total": out["mixed_multigrid"]["n_cg"],
            "multigrid_compliance_rel_err_final": float(abs(out["mixed_multigrid"]["c"][-1] - out["fp64"]["c"][-1])
                                                        / abs(out["fp64"]["c"][-1])),
            "compliance_rel_err_final": float(rel[-1]), "compliance_rel_err_max": float(rel.max()),
            "pcg_gbs_algorithmic": round(b_it * m["n_cg"] / m["seconds"] / 1e9, 1),
            "dtype": "f32 vectors + f64 dots/refresh/design fields", "data": "synthetic"}


def serving_bench(log, n=512, rate=20.0, max_active=64, seq_sample=32, mode="auto", preconditioner=MAIN_PC):
    """Serving stream: n seeded requests, Poisson arrivals at `rate`/s, nelx
    uniform over multiples of 16 in [64, 768], aspect 1/4, 1/3 or 1/2, BC
    cantilever / MBB / bridge, volfrac U[0.3, 0.5], rmin 0.008 nelx + 1, run to
    100 OC iterations or max change < 0.01; continuous batching with HBM
    admission (mode 'async': every design runs its own SIMP iterations;
    'lockstep': one graph-replayed iteration of every active design per
    engine step; 'auto': lockstep for multigrid designs, async otherwise).
    Sequential baseline: the first seq_sample requests of the same stream
    served one at a time (designs/hour of that sample)."""
    from toptimization_amd.serving import DesignServer, synthetic_stream
    solver = dict(cg_tolerance=1e-6, max_cg_iterations=50000, preconditioner=preconditioner,
                  precision="mixed")
    reqs = synthetic_stream(n=n, rate=rate, seed=0, nelx_range=(64, 768), max_iter=100, tol=0.01, **solver)
    server = DesignServer(max_active=max_active, hbm_budget=64 << 30, precision="mixed", mode=mode)
    st = server.run(reqs, clock="wall", progress=log)
    s = st.summary()
    log(f"  serving {n} requests at {rate}/s: makespan {s['makespan_s']:.1f} s, "
        f"{s['designs_per_hour']:.0f} designs/h, p50 {s['p50_latency_s']:.2f} s, p99 {s['p99_latency_s']:.2f} s, "
        f"max active {s['max_active']}")
    seq = None
    if seq_sample:
        sub = synthetic_stream(n=seq_sample, rate=rate, seed=0, nelx_range=(64, 768), max_iter=100,
                               tol=0.01, **solver)
        one = DesignServer(max_active=1, hbm_budget=64 << 30, precision="mixed")
        ss = one.run(sub, clock="instant", progress=log).summary()
        seq = ss["designs_per_hour"]
        log(f"  sequential, first {seq_sample} requests back to back: {seq:.0f} designs/h")
    return {"metric": f"serving designs/hour ({n}-request Poisson stream)", "value": round(s["designs_per_hour"], 1),
            "unit": "designs/hour", "p50_latency_s": round(s["p50_latency_s"], 3),
            "p99_latency_s": round(s["p99_latency_s"], 3), "makespan_s": round(s["makespan_s"], 2),
            "requests": n, "arrival_rate_per_s": rate, "max_active": s["max_active"], "mode": mode,
            "preconditioner": preconditioner,
            "engine_steps": s["engine_steps"], "sequential_designs_per_hour": None if seq is None else round(seq, 1),
            "sequential_sample": seq_sample,
            "sequential_note": (f"the first {seq_sample} requests of the same stream served one at a time "
                                "(max_active 1), designs/hour of that sample"),
            "dtype": "f32 vectors + f64 dots/refresh", "data": "synthetic"}


def h8_bench(log, nelx=96, nely=48, nelz=48, iters=6, reps=50):
    """3-D H8 path: a 3-D cantilever (left face clamped, unit downward load
    at the middle of the right face's bottom edge), volfrac 0.3, rmin 1.5,
    FP64 Jacobi PCG rtol 1e-6 on the device; OC it/s over `iters` OC
    iterations after one warm-up, CG it/s, and the FP64 3-D matvec (read p,
    write Ap: 24 B per node each, read s: 8 B per element) with HIP events
    over back-to-back launches."""
    import numpy as np
    import torch
    from toptimization_amd import h8
    from toptimization_amd.problem import make_problem
    prob = make_problem(nelx, nely, nelz=nelz,
                        fixed=[{"nodes": {"type": "face", "face": "left"}, "dofs": ["x", "y", "z"]}],
                        loads=[{"nodes": {"type": "point", "location": [nelx, 0, nelz // 2]},
                                "force": [0.0, -1.0, 0.0]}],
                        name="cantilever3d", volume_fraction=0.3, filter_radius=1.5, max_iterations=iters + 1,
                        tolerance=0.0, cg_tolerance=1e-6, max_cg_iterations=50000, preconditioner="jacobi")
    d = h8.Design3(prob)
    st = d.iteration()
    log(f"  [h8 {nelx}x{nely}x{nelz}] warm-up OC iteration: c={st['c']:.6e} n_cg={st['n_cg']}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ncg = 0
    for _ in range(iters):
        st = d.iteration()
        ncg += st["n_cg"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"  [h8] {iters} OC iterations in {dt:.3f} s ({iters / dt:.2f} OC it/s), {ncg} CG iterations "
        f"({ncg / dt:.0f} CG it/s)")
    g = d.g
    p = torch.as_tensor(np.random.default_rng(1).normal(size=g.n_dofs), device="cuda")
    Ap = torch.empty_like(p)
    for _ in range(3):
        h8.matvec(g, p, d.s, d.fix, Ap=Ap)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        h8.matvec(g, p, d.s, d.fix, Ap=Ap)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    b = 48 * g.n_nodes + 8 * g.n_elem
    log(f"  [h8] matvec {ms * 1e3:.1f} us, {g.n_dofs / (ms * 1e-3) / 1e9:.2f} GDOF/s, "
        f"{b / (ms * 1e-3) / 1e9:.0f} GB/s algorithmic")
    return {"metric": f"OC it/s (3-D H8 cantilever {nelx}x{nely}x{nelz}, FP64 Jacobi PCG rtol 1e-6)",
            "value": round(iters / dt, 3), "unit": "OC it/s", "cg_its_per_s": round(ncg / dt, 1),
            "cg_per_oc_it": round(ncg / iters, 1), "dofs": g.n_dofs, "elements": g.n_elem,
            "matvec_us": round(ms * 1e3, 2), "matvec_gdofs": round(g.n_dofs / (ms * 1e-3) / 1e9, 2),
            "matvec_gbs_algorithmic": round(b / (ms * 1e-3) / 1e9, 1), "dtype": "f64", "data": "synthetic"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="headline only: skip the config 1, 3 and 4 measurements of the default line")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="profiling driver: three multigrid OC iterations, one block-Jacobi OC iteration, "
                         "100 slot-A launches, op-level matvecs, copy calibration")
    ap.add_argument("--batch", action="store_true",
                    help="64 x 128x64 cantilevers: one batched plan vs one at a time (own JSON line)")
    ap.add_argument("--serving", action="store_true",
                    help="512-request serving stream: designs/hour, p50/p99 latency (own JSON line)")
    ap.add_argument("--mbb", action="store_true",
                    help="300x100 MBB, 100 OC iterations: mixed vs FP64 vs Jacobi (own JSON line)")
    ap.add_argument("--h8", action="store_true",
                    help="3-D H8 path: 96x48x48 cantilever OC it/s and 3-D matvec (own JSON line)")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--serving-mode", choices=["auto", "async", "lockstep"], default="auto")
    ap.add_argument("--pc", default=None, help="preconditioner of the --batch / --serving runs "
                    "(default: multigrid)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); a design problem is "
              "solved on one device and independent problems shard by problem, so there is no "
              "multi-GPU figure to report", file=sys.stderr)
        return 2
    log = (lambda *a: None) if args.quiet else (lambda *a: print(*a, file=sys.stderr, flush=True))
    import torch
    torch.cuda.set_device(0)
    nn, ne = (NELX + 1) * (NELY + 1), NELX * NELY
    if args.batch:
        print(json.dumps(batch_bench(log, preconditioner=args.pc or MAIN_PC)), flush=True)
        return 0
    if args.mbb:
        print(json.dumps(mbb_bench(log)), flush=True)
        return 0
    if args.h8:
        print(json.dumps(h8_bench(log)), flush=True)
        return 0
    if args.serving:
        print(json.dumps(serving_bench(log, n=args.requests, seq_sample=args.seq_sample, mode=args.serving_mode,
                                         preconditioner=args.pc or MAIN_PC)), flush=True)
        return 0
    if args.probe:
        from toptimization_amd.engine import DesignState, Plan
        st = DesignState(bridge(NELX, NELY, "mixed", preconditioner="multigrid"))
        plan = Plan([st])
        for k in range(3):
            s = plan.iteration()[0]
            log(f"[probe] multigrid OC iteration {k + 1}: c={s.compliance:.6e} n_cg={s.n_cg}")
        plan.close()
        st = DesignState(bridge(NELX, NELY, "mixed"))
        plan = Plan([st])
        s = plan.iteration()[0]
        log(f"[probe] 2x2 block Jacobi OC iteration 1: c={s.compliance:.6e} n_cg={s.n_cg}")
        kernel_roofline(plan, nn, ne, log, reps=100)
        matvec_gdofs(NELX, NELY, log, reps=50)
        matvec_gdofs(8192, 4096, log, reps=10)
        matvec_gdofs(8192, 4096, log, reps=10, dtype="f64")
        plan.close()
        # calibration for HBM byte counters: a plain device copy of 256 MiB
        # (reads and writes exactly 256 MiB each), three times
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[probe] copy calibration: 3 x 256 MiB device copy")
        return 0

    log(f"[bench] {NELX}x{NELY} bridge, warmup {args.warmup}, steps {args.steps}")
    mixed = run_path("mixed", args.warmup, args.steps, log)
    ms_a, bytes_a, gbs_a = kernel_roofline(mixed["plan"], nn, ne, log)
    rho_probe = mixed["state"].rho.cpu().numpy()
    mixed["plan"].close()
    full = full_run(log)
    fp64 = run_path("fp64", args.warmup, args.steps, log)
    fp64["plan"].close()
    # the same workload with the nodal 2x2 block-Jacobi preconditioner (a
    # few iterations: ~4,800 CG iterations each)
    bj = run_path("mixed", 1, min(args.steps, 5), log, preconditioner="block_jacobi_2x2")
    bj["plan"].close()
    rel_c = same_rho_compliance(rho_probe, log)
    mv_ms, mv_gdofs, mv_gbs = matvec_gdofs(NELX, NELY, log)
    mvb_ms, mvb_gdofs, mvb_gbs = matvec_gdofs(8192, 4096, log, reps=20)
    _, mvd_gdofs, mvd_gbs = matvec_gdofs(8192, 4096, log, reps=10, dtype="f64")
    cpu = None if args.no_cpu else cpu_baseline(log)
    # the other benchmark configurations, measured in the same run (their
    # own sub-objects; the headline value above is config 2's)
    configs = None
    if not args.quick:
        configs = {"mbb_300x100": mbb_bench(log),
                   "batch_64x128x64": batch_bench(log, seq_sample=8),
                   "serving_512": serving_bench(log, n=512, seq_sample=32)}

    value = args.steps / mixed["seconds"]
    fp64_value = args.steps / fp64["seconds"]
    # roofline of the headline's unit of work: one multigrid CG iteration
    # (11 launches), algorithmic bytes over device time per CG iteration
    tm = mixed["timing"]
    us_cg = 1e3 * tm["pcg"] / max(tm["n_cg"], 1)
    b_cg = mg_cg_bytes(NELX, NELY)
    gbs_cg = b_cg / (us_cg * 1e-6) / 1e9
    log(f"  multigrid CG iteration: {b_cg / 1e6:.2f} MB algorithmic in {us_cg:.1f} us -> {gbs_cg:.0f} GB/s")
    line = {
        "metric": "FEM matvec GDOF/s, OC iters/s, serving p50/p99 latency at compliance err <=1e-6",
        "value": round(value, 3),
        "unit": "OC it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 vectors + f64 dots/refresh/design fields",
        "data": "synthetic (bridge load case, rho0 = volfrac)",
        "config": {"workload": f"SIMP bridge {NELX}x{NELY} Q4, volfrac 0.3, rmin 6, p 3, "
                               "PCG rtol 1e-6 mixed FP32/FP64, geometric multigrid V(1,1) preconditioner",
                   "nelx": NELX, "nely": NELY, "dofs": 2 * nn, "elements": ne,
                   "parallelism": "single-gpu"},
        "pcg_its_per_s": round(mixed["n_cg"] / mixed["seconds"], 1),
        "cg_per_oc_it": round(mixed["n_cg"] / args.steps, 1),
        "fp64_oc_its_per_s": round(fp64_value, 3),
        "fp64_cg_per_oc_it": round(fp64["n_cg"] / args.steps, 1),
        "mixed_over_fp64": round(value / fp64_value, 3),
        "us_per_cg_it": round(1e6 * mixed["seconds"] / max(mixed["n_cg"], 1), 1),
        "block_jacobi_2x2_oc_its_per_s": round(len(bj["stats"]) / bj["seconds"], 3),
        "block_jacobi_2x2_cg_per_oc_it": round(bj["n_cg"] / len(bj["stats"]), 1),
        "block_jacobi_2x2_oc_iterations": f"2-{1 + len(bj['stats'])}",
        "multigrid_over_block_jacobi_2x2": round(value / (len(bj["stats"]) / bj["seconds"]), 2),
        "same_rho_compliance_rel_err": rel_c,
        "matvec_gdofs": round(mv_gdofs, 2),
        "matvec_gbs_algorithmic": round(mv_gbs, 1),
        "matvec_8192x4096_gdofs": round(mvb_gdofs, 2),
        "matvec_8192x4096_gbs": round(mvb_gbs, 1),
        "matvec_f64_8192x4096_gdofs": round(mvd_gdofs, 2),
        "matvec_f64_8192x4096_gbs": round(mvd_gbs, 1),
        "roofline": {"bound": "hbm",
                     "kernel": "multigrid PCG iteration, 11 launches: slot A; slot B fused with the level-0 "
                               "down sweep and restriction; down sweeps 1, 2 and 3+4 (one launch), tail "
                               "(levels 5-7), up sweeps 4-1; slot C fused with the level-0 up sweep",
                     "achieved": round(gbs_cg, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs_cg / HBM_PEAK_GBS, 4),
                     "traffic": profile_traffic(PROFILE_MG, "cg_iteration"),
                     "bytes_per_cg_iteration": b_cg, "us_per_cg_iteration": round(us_cg, 2),
                     "timing": "HIP events on the launch stream around the PCG chunks of the timed OC "
                               "iterations, replayed from a fresh state (same CG count)"},
        "phase_ms_per_oc_iteration": {k: round(tm[k] / args.steps, 4) for k in ("setup", "pcg", "post")},
        "matvec_roofline": {"kernel": "k_matvec_modal<float, 7> (op-level K(s) p, 8192x4096, FP32)",
                            "achieved": round(mvb_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(mvb_gbs / HBM_PEAK_GBS, 4),
                            "traffic": profile_traffic(PROFILE_MV, "matvec_f32_8192x4096"),
                            "bytes_per_launch": 4 * (4 * 8193 * 4097 + 8192 * 4096),
                            "us_per_launch": round(mvb_ms * 1e3, 2)},
        "slot_a_roofline": {"kernel": "k_slot_a<float> (PCG slot A)", "achieved": round(gbs_a, 1),
                            "frac": round(gbs_a / HBM_PEAK_GBS, 4),
                            "traffic": profile_traffic(PROFILE_SLOTS, "slot_a_launch"),
                            "bytes_per_launch": bytes_a, "us_per_launch": round(ms_a * 1e3, 3)},
        "full_run_200": {"seconds": round(full["seconds"], 3), "oc_its_per_s": round(200 / full["seconds"], 2),
                         "cg_iterations": full["n_cg"], "final_compliance": full["final_compliance"]},
        "cpu_baseline": cpu,
        "configs": configs,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
