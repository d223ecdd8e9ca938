"""SIMP topology optimization main loop on the GPU engine.

Drop-in for the reference ``optimizer/simp.py:19-231`` (``run``): same
signature, same per-iteration printout, same outputs
(``density_final.npy`` flat float32 in ``e = ex*nely+ey`` order,
``density_final.png``, ``convergence.png``), same stopping rule
(``it > 5 and change < tol``, ``simp.py:217``). One loop iteration is one
``topt_plan_iteration`` call: filter -> PCG -> sensitivity -> dc filter ->
OC, with a single device->host read of the iteration scalars.
"""

from __future__ import annotations

import time
from pathlib import Path
from typing import Callable

import numpy as np

from ..engine import DesignState, Plan


def run(problem, backend: str = "auto", live_viz: bool = True,
        viz_callback: Callable | None = None, *, verbose: bool = True,
        history: dict | None = None) -> np.ndarray:
    """Run SIMP for ``problem``; returns the final density as (nelx, nely).

    ``backend``: "auto" or "hip" (the only compute path; any other value
    raises, since the reference's Taichi backends do not exist here).
    ``history`` (optional dict) receives per-iteration lists
    ``c, vol, change, n_cg, oc_steps, time``.
    """
    if backend not in ("auto", "hip"):
        raise ValueError(f"Unknown backend '{backend}': this build runs on HIP (gfx950) only")
    if problem.dim == 3:
        # the 3-D H8 path: op-level kernels and the device PCG (h8.py)
        from ..h8 import run_3d
        return run_3d(problem, verbose=verbose, history=history)
    log = print if verbose else (lambda *a, **k: None)
    log("[toptimization] Backend: hip (gfx950)")
    log(f"[toptimization] Building mesh ({problem.nelx}x{problem.nely})...")
    log(f"[toptimization] Allocating fields (n_dofs={problem.n_dofs:,}, n_elem={problem.n_elem:,})...")
    state = DesignState(problem)
    plan = Plan([state])
    out_dir = Path(problem.output_dir)
    out_dir.mkdir(parents=True, exist_ok=True)

    axes = None
    if live_viz:
        try:
            import matplotlib
            import matplotlib.pyplot as plt
            plt.ion()
            fig, axes = plt.subplots(1, 2, figsize=(12, 4))
            fig.tight_layout()
        except Exception:
            live_viz = False

    log(f"\n[toptimization] Starting: {problem.name}")
    log(f"{'Iter':>5} {'Compliance':>14} {'Volume':>8} {'Change':>9} {'CG':>6} {'s/iter':>7}")
    log("-" * 55)
    hist = history if history is not None else {}
    for k in ("c", "vol", "change", "n_cg", "oc_steps", "time"):
        hist.setdefault(k, [])
    sens_mode = problem.filter_type == "sensitivity"
    for it in range(1, problem.max_iter + 1):
        t0 = time.perf_counter()
        st = plan.iteration(sensitivity_filter=sens_mode)[0]
        elapsed = time.perf_counter() - t0
        for key, val in (("c", st.compliance), ("vol", st.volume), ("change", st.change),
                         ("n_cg", st.n_cg), ("oc_steps", st.oc_steps), ("time", elapsed)):
            hist[key].append(val)
        log(f"{it:>5d} {st.compliance:>14.6f} {st.volume:>8.4f} {st.change:>9.4f} "
            f"{st.n_cg:>6d} {elapsed:>7.2f}")
        need_rho = (live_viz and axes is not None) or viz_callback is not None or \
            it % problem.save_interval == 0
        if need_rho:
            rho_np = state.rho.double().cpu().numpy()
            if live_viz and axes is not None and it % max(1, problem.save_interval // 5) == 0:
                _update_live_fig(axes, rho_np, problem, it, hist["c"])
            if it % problem.save_interval == 0:
                _save_iter_png(problem, rho_np, it)
            if viz_callback is not None:
                viz_callback(rho_np, it, st.compliance, st.volume)
        if it > 5 and st.change < problem.tol:
            log(f"\n[toptimization] Converged at iter {it} (change={st.change:.6f})")
            break
    log("\n[toptimization] Done.")
    rho_np = state.rho.cpu().numpy()
    hist["u"] = state.u.cpu().numpy()
    _save_final(problem, rho_np, hist["c"], hist["vol"], log)
    plan.close()
    if live_viz:
        try:
            import matplotlib.pyplot as plt
            plt.ioff()
            plt.show(block=True)
        except Exception:
            pass
    return rho_np.reshape(problem.nelx, problem.nely)


def _update_live_fig(axes, rho_np, problem, it, compliance_hist):
    import matplotlib.pyplot as plt
    axes[0].cla()
    axes[0].imshow(1.0 - rho_np.reshape(problem.nelx, problem.nely).T, cmap="gray", vmin=0, vmax=1,
                   origin="lower", aspect="equal")
    axes[0].set_title(f"iter {it}")
    axes[0].axis("off")
    axes[1].cla()
    axes[1].semilogy(compliance_hist, linewidth=1.2)
    axes[1].set_xlabel("Iteration")
    axes[1].set_ylabel("Compliance")
    plt.pause(0.001)


def _save_iter_png(problem, rho_np, it):
    try:
        from ..viz import save_density_png
        save_density_png(rho_np.reshape(problem.nelx, problem.nely),
                         Path(problem.output_dir) / f"iter_{it:04d}.png",
                         title=f"{problem.name} — iter {it}")
    except Exception:
        pass


def _save_final(problem, rho_np, c_hist, v_hist, log):
    """``simp.py:282-306``: density_final.npy (float32, flat) + PNGs."""
    out = Path(problem.output_dir)
    np.save(out / "density_final.npy", rho_np.astype(np.float32))
    try:
        from ..viz import save_convergence_png, save_density_png
        save_density_png(rho_np.reshape(problem.nelx, problem.nely), out / "density_final.png",
                         title=problem.name)
        save_convergence_png(c_hist, v_hist, out / "convergence.png")
        log(f"[toptimization] Saved results to {out}/")
    except Exception as e:  # matplotlib missing or headless failure
        log(f"[warn] Save failed: {e}")
