"""Command line: ``toptimization run|validate <problem.yaml>`` (reference ``cli.py:22-144``).

``--backend`` accepts auto|hip (the GPU engine is the only compute path).
"""

from __future__ import annotations

import argparse
import sys
from pathlib import Path


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="toptimization",
                                 description="SIMP topology optimization on AMD MI355X (HIP)")
    sub = ap.add_subparsers(dest="command")
    run_p = sub.add_parser("run", help="Run topology optimization")
    run_p.add_argument("problem")
    run_p.add_argument("--backend", choices=["auto", "hip"], default="auto")
    run_p.add_argument("--no-viz", action="store_true")
    run_p.add_argument("--no-live", action="store_true")
    run_p.add_argument("--output", metavar="DIR")
    run_p.add_argument("--show", action="store_true")
    run_p.add_argument("--precision", choices=["fp64", "mixed"])
    val_p = sub.add_parser("validate", help="Validate a problem file without running")
    val_p.add_argument("problem")
    srv_p = sub.add_parser("serve", help="Serve a JSON-lines file of design requests (continuous batching)")
    srv_p.add_argument("requests", help="one JSON request per line: {id, arrival, problem: {...}}")
    srv_p.add_argument("--output", metavar="DIR", default="serve_results")
    srv_p.add_argument("--max-active", type=int, default=64)
    srv_p.add_argument("--hbm-budget-gb", type=float, default=64.0)
    srv_p.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    srv_p.add_argument("--clock", choices=["wall", "instant"], default="instant",
                       help="wall: honour arrival times; instant: all requests available at t=0")
    srv_p.add_argument("--mode", choices=["auto", "async", "lockstep"], default="auto",
                       help="async: every design runs its own SIMP iterations; lockstep: one "
                            "graph-replayed iteration of every active design per step; auto: "
                            "lockstep for multigrid designs, async for the others")
    args = ap.parse_args(argv)
    if args.command == "run":
        return _cmd_run(args)
    if args.command == "serve":
        return _cmd_serve(args)
    if args.command == "validate":
        return _cmd_validate(args)
    ap.print_help()
    return 0


def _cmd_validate(args) -> int:
    from .problem import load_problem
    path = Path(args.problem)
    if not path.exists():
        print(f"Error: file not found: {path}", file=sys.stderr)
        return 1
    try:
        p = load_problem(path)
    except Exception as e:
        print(f"Validation failed: {e}", file=sys.stderr)
        return 1
    print(f"Problem '{p.name}' is valid.")
    print(f"  Dimension   : {p.dim}D")
    print(f"  Mesh        : {p.nelx} x {p.nely}" + (f" x {p.nelz}" if p.dim == 3 else ""))
    print(f"  Elements    : {p.n_elem:,}")
    print(f"  DOFs        : {p.n_dofs:,}")
    print(f"  Fixed DOFs  : {len(p.fixed_dofs):,}")
    print(f"  Force DOFs  : {len(p.force_dofs):,}")
    print(f"  Volume frac : {p.volfrac}")
    print(f"  Filter      : {p.filter_type} (rmin={p.rmin})")
    print(f"  Max iter    : {p.max_iter}")
    print(f"  Solver      : {p.preconditioner}, {p.precision}")
    return 0


def _cmd_run(args) -> int:
    from .optimizer.simp import run
    from .problem import load_problem
    path = Path(args.problem)
    if not path.exists():
        print(f"Error: file not found: {path}", file=sys.stderr)
        return 1
    try:
        problem = load_problem(path)
    except Exception as e:
        print(f"Error loading problem: {e}", file=sys.stderr)
        return 1
    if args.output:
        problem.output_dir = Path(args.output)
    if args.precision:
        problem.precision = args.precision
    try:
        rho = run(problem=problem, backend=args.backend, live_viz=not (args.no_viz or args.no_live))
    except Exception as e:
        import traceback
        print(f"Error during optimization: {e}", file=sys.stderr)
        traceback.print_exc()
        return 1
    if args.show and not args.no_viz:
        try:
            import matplotlib.pyplot as plt
            plt.imshow(1.0 - rho.T, cmap="gray", origin="lower")
            plt.show(block=True)
        except Exception as e:
            print(f"[warn] Could not show result: {e}")
    print(f"\nResults saved to: {problem.output_dir}/")
    return 0


def _cmd_serve(args) -> int:
    import json

    import numpy as np

    from .serving import DesignServer, load_requests
    try:
        reqs = load_requests(args.requests)
    except Exception as e:
        print(f"Error reading requests: {e}", file=sys.stderr)
        return 1
    out = Path(args.output)
    out.mkdir(parents=True, exist_ok=True)
    server = DesignServer(max_active=args.max_active, hbm_budget=int(args.hbm_budget_gb * (1 << 30)),
                          precision=args.precision, keep_rho=True, mode=args.mode)
    stats = server.run(reqs, clock=args.clock, progress=lambda m: print(m, flush=True))
    by_id = {r.rid: r for r in reqs}
    rows = []
    for res in stats.results:
        p = by_id[res.rid].problem
        np.save(out / f"{res.rid}_density.npy", res.rho.astype(np.float32).reshape(p.nelx, p.nely))
        rows.append({"id": res.rid, "nelx": p.nelx, "nely": p.nely, "iterations": res.iterations,
                     "compliance": res.compliance, "volume": res.volume, "change": res.change,
                     "cg_iterations": res.n_cg, "arrival_s": res.arrival, "finished_s": res.finished,
                     "latency_s": res.latency})
    summary = {"summary": stats.summary(), "designs": rows}
    (out / "summary.json").write_text(json.dumps(summary, indent=1))
    print(json.dumps(stats.summary()))
    print(f"Results saved to: {out}/")
    return 0


if __name__ == "__main__":
    sys.exit(main())
