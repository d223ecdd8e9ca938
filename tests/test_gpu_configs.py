"""The workloads of the benchmark configurations at their own sizes, the
sensitivity-filter loop, the result files of ``toptimization run``, and the
drop-in driven through the reference's own import name (``toptimization``).

Checkers: the FP64 NumPy oracle (small grids) and its C restatement (large
grids), both pinned in tests/test_oracle.py."""

import numpy as np
import pytest

from oracle import simp_np as O

from helpers import bridge, cantilever, oracle_args

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b) / np.abs(b)


def _c_ref(prob, cg_tol=1e-10):
    """The C restatement in FP64 with converged Jacobi solves (same problem,
    same OC loop and stopping rule)."""
    import dataclasses

    from oracle import c_oracle
    c_oracle.set_threads(16)
    return c_oracle.simp(dataclasses.replace(prob, cg_tol=cg_tol, max_cg_iter=200000,
                                             preconditioner="jacobi", precision="fp64"), "f64")


# ---------------------------------------------------------------------------
# config 0: 60x20 cantilever, displacements (optimizer/simp.py:141-219)
# ---------------------------------------------------------------------------


def test_config0_displacements_match_oracle():
    """Final u of the 50-iteration FP64 Jacobi loop (rtol 1e-8) against the
    oracle's; compliance is checked per iteration in test_gpu_parity."""
    from toptimization_amd.optimizer.simp import run
    prob = cantilever(60, 20, volume_fraction=0.5, filter_radius=1.5, max_iterations=50,
                      tolerance=0.0, cg_tolerance=1e-8, max_cg_iterations=20000)
    args, kw = oracle_args(prob)
    ref = O.simp_run(*args, **kw)
    hist = {}
    run(prob, live_viz=False, verbose=False, history=hist)
    u, uref = hist["u"], ref["u"]
    err = np.abs(u - uref).max() / np.abs(uref).max()
    print(f"config 0: max |u - u_ref| / max |u_ref| = {err:.2e}", flush=True)
    # both solves stop at |r| < 1e-8 |f|; K's condition number bounds the
    # displacement error well inside this
    assert err <= 1e-6
    free = np.ones(prob.n_dofs, bool)
    free[np.asarray(prob.fixed_dofs)] = False
    assert np.all(u[~free] == 0.0)


# ---------------------------------------------------------------------------
# sensitivity-filter mode (optimizer/simp.py:174-189, filter.py:152-171)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("precision,pc", [("fp64", "jacobi"), ("mixed", "multigrid")])
def test_sensitivity_filter_loop_matches_oracle(precision, pc):
    from toptimization_amd.optimizer.simp import run
    prob = cantilever(60, 20, volume_fraction=0.5, filter_radius=1.5, filter_type="sensitivity",
                      max_iterations=30, tolerance=0.0, cg_tolerance=1e-8, max_cg_iterations=20000,
                      precision=precision, preconditioner=pc)
    args, kw = oracle_args(prob)
    kw["preconditioner"] = "jacobi"
    ref = O.simp_run(*args, **kw)
    hist = {}
    rho = run(prob, live_viz=False, verbose=False, history=hist)
    rel = _rel(hist["c"], ref["c"])
    drho = np.abs(rho.ravel() - ref["rho"]).max()
    print(f"sensitivity mode {precision}/{pc}: max rel c err {rel.max():.2e}, max |rho err| {drho:.2e}",
          flush=True)
    assert len(hist["c"]) == 30
    assert rel.max() <= 1e-6
    np.testing.assert_allclose(hist["vol"], ref["vol"], rtol=1e-8)
    assert drho <= 1e-5


# ---------------------------------------------------------------------------
# config 2: 1200x400 bridge, mixed FP32/FP64 against plain FP64 along the
# OC trajectory, and OC iteration 1 against the C restatement in FP64
# ---------------------------------------------------------------------------


def test_config2_mixed_trajectory_matches_fp64():
    from toptimization_amd.engine import DesignState, Plan
    n_it = 20
    c = {}
    for prec in ("mixed", "fp64"):
        prob = bridge(1200, 400, volume_fraction=0.3, filter_radius=6.0, max_iterations=200, tolerance=0.0,
                      cg_tolerance=1e-6, max_cg_iterations=50000, preconditioner="multigrid", precision=prec)
        st = DesignState(prob)
        plan = Plan([st])
        c[prec] = [plan.iteration()[0].compliance for _ in range(n_it)]
        plan.close()
        print(f"config 2 {prec}: c[1]={c[prec][0]:.9e} c[{n_it}]={c[prec][-1]:.9e}", flush=True)
    rel = _rel(c["mixed"], c["fp64"])
    print(f"config 2: mixed vs fp64 max rel compliance difference over {n_it} OC iterations {rel.max():.2e}",
          flush=True)
    assert rel.max() <= 1e-6


@pytest.mark.timeout(600)
def test_config2_first_iteration_matches_c_oracle():
    from toptimization_amd.engine import DesignState, Plan
    prob = bridge(1200, 400, volume_fraction=0.3, filter_radius=6.0, max_iterations=1, tolerance=0.0,
                  cg_tolerance=1e-6, max_cg_iterations=50000, preconditioner="multigrid", precision="mixed")
    st = DesignState(prob)
    plan = Plan([st])
    s = plan.iteration()[0]
    rho = st.rho.cpu().numpy()
    plan.close()
    ref = _c_ref(prob, cg_tol=1e-8)
    rel = abs(s.compliance - ref["c"][0]) / ref["c"][0]
    drho = np.abs(rho - ref["rho"]).max()
    print(f"config 2 OC iteration 1: GPU mixed multigrid c={s.compliance:.12e}, C oracle FP64 "
          f"c={ref['c'][0]:.12e} ({ref['n_cg'][0]} CG, {ref['seconds']:.1f} s): rel {rel:.2e}, "
          f"max |rho err| {drho:.2e}", flush=True)
    assert rel <= 1e-6
    assert drho <= 1e-5
    assert abs(s.volume - ref["vol"][0]) <= 1e-9


# ---------------------------------------------------------------------------
# config 3: 64 cantilevers of 128x64 in one ragged plan
# ---------------------------------------------------------------------------


@pytest.mark.timeout(600)
def test_config3_batch_of_64_at_size():
    """All 64 problems of the batch (mixed FP32/FP64, multigrid, rtol 1e-6)
    bitwise equal to their solo runs over 3 OC iterations; two members
    against the oracle in FP64."""
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import batch_cantilevers
    n_it = 3
    probs = batch_cantilevers(64, max_iter=n_it, cg_tolerance=1e-6, max_cg_iterations=50000,
                              preconditioner="multigrid", precision="mixed")
    states = [DesignState(p) for p in probs]
    plan = Plan(states)
    mix = [plan.iteration() for _ in range(n_it)]
    plan.close()
    print(f"config 3: batched plan of {len(probs)} x 128x64 done, CG per problem at iteration 1: "
          f"{min(s.n_cg for s in mix[0])}-{max(s.n_cg for s in mix[0])}", flush=True)
    for i, pr in enumerate(probs):
        st = DesignState(pr)
        pl = Plan([st])
        alone = [pl.iteration()[0] for _ in range(n_it)]
        pl.close()
        for k in range(n_it):
            assert (mix[k][i].compliance, mix[k][i].n_cg, mix[k][i].change) == \
                (alone[k].compliance, alone[k].n_cg, alone[k].change), (i, k)
        np.testing.assert_array_equal(states[i].rho.cpu().numpy(), st.rho.cpu().numpy())
    for i in (0, 37):
        ref = _c_ref(probs[i])
        got = [m[i].compliance for m in mix]
        rel = _rel(got, ref["c"])
        drho = np.abs(states[i].rho.cpu().numpy() - ref["rho"]).max()
        print(f"config 3 member {i}: max rel c err vs oracle {rel.max():.2e}, max |rho err| {drho:.2e}",
              flush=True)
        assert rel.max() <= 1e-6
        assert drho <= 1e-5


# ---------------------------------------------------------------------------
# config 4: the serving stream (64..768 wide designs, three BC types)
# ---------------------------------------------------------------------------


@pytest.mark.timeout(600)
def test_config4_stream_prefix_equals_solo_and_oracle():
    """First 24 requests of the config-4 stream through the continuous-
    batching server (8 slots): every design bitwise equal to its solo run;
    the two smallest against the FP64 oracle along their whole runs."""
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import DesignServer, synthetic_stream
    solver = dict(cg_tolerance=1e-6, max_cg_iterations=50000, preconditioner="multigrid", precision="mixed")
    reqs = synthetic_stream(n=24, rate=20.0, seed=0, nelx_range=(64, 768), max_iter=100, tol=0.01, **solver)
    server = DesignServer(max_active=8, hbm_budget=16 << 30, precision="mixed", keep_rho=True)
    stats = server.run(reqs, clock="instant")
    s = stats.summary()
    print(f"config 4 prefix: {s['designs']} designs, makespan {s['makespan_s']:.2f} s, max active "
          f"{s['max_active']}", flush=True)
    assert s["designs"] == 24 and s["max_active"] == 8
    hists = {}
    for res in stats.results:
        prob = reqs[res.rid].problem
        st = DesignState(prob, precision="mixed")
        plan = Plan([st])
        h = []
        for _ in range(res.iterations):
            h.append(plan.iteration()[0])
        plan.close()
        hists[res.rid] = h
        assert (h[-1].compliance, h[-1].volume, h[-1].change) == (res.compliance, res.volume, res.change), res.rid
        np.testing.assert_array_equal(st.rho.cpu().numpy(), res.rho)
        # the server stopped it where the reference's rule stops it
        stop = next(k + 1 for k, x in enumerate(h) if (k + 1 > 5 and x.change < prob.tol) or k + 1 == prob.max_iter)
        assert stop == res.iterations, res.rid
    smallest = sorted(stats.results, key=lambda r: reqs[r.rid].problem.n_elem)[:2]
    for res in smallest:
        prob = reqs[res.rid].problem
        ref = _c_ref(prob)
        got = [x.compliance for x in hists[res.rid]]
        n = min(len(got), len(ref["c"]))
        rel = _rel(got[:n], ref["c"][:n])
        print(f"config 4 request {res.rid} ({prob.name}, {prob.nelx}x{prob.nely}): {len(got)} GPU vs "
              f"{len(ref['c'])} oracle iterations, max rel c err {rel.max():.2e}", flush=True)
        assert rel.max() <= 5e-5
        assert len(got) == len(ref["c"])
        assert rel[-1] <= 1e-6


# ---------------------------------------------------------------------------
# toptimization run: result files (cli.py:103-144, optimizer/simp.py:282-306)
# ---------------------------------------------------------------------------

SMALL_CANTILEVER_YAML = """\
problem:
  name: "Cantilever Small Test"
  dimension: 2
domain:
  nelx: 20
  nely: 10
material:
  E: 1.0
  nu: 0.3
  E_min: 1.0e-9
boundary_conditions:
  fixed:
    - nodes:
        type: "edge"
        edge: "left"
      dofs: ["x", "y"]
loads:
  - nodes:
      type: "point"
      location: [20, 5]
    force: [0.0, -1.0]
optimization:
  volume_fraction: 0.5
  penalty: 3.0
  filter_radius: 1.5
  filter_type: "density"
  max_iterations: 50
  tolerance: 0.01
  move_limit: 0.2
solver:
  max_cg_iterations: 500
  cg_tolerance: 1.0e-8
  preconditioner: "jacobi"
output:
  directory: "results/cantilever_small"
  save_interval: 10
  format: "png"
"""


def test_cli_run_writes_density_final(tmp_path):
    """``toptimization run problem.yaml --no-viz --output DIR`` (through the
    reference's module name): density_final.npy is flat float32 in
    e = ex*nely + ey order, equal to the oracle's final density."""
    import toptimization.cli as cli
    from toptimization.problem import load_problem
    y = tmp_path / "cantilever_small.yaml"
    y.write_text(SMALL_CANTILEVER_YAML)
    out = tmp_path / "out"
    assert cli.main(["run", str(y), "--no-viz", "--output", str(out)]) == 0
    rho = np.load(out / "density_final.npy")
    assert rho.dtype == np.float32 and rho.shape == (200,)
    prob = load_problem(y)
    args, kw = oracle_args(prob)
    ref = O.simp_run(*args, **kw)
    err = np.abs(rho.astype(np.float64) - ref["rho"]).max()
    print(f"cli run: {len(ref['c'])} oracle iterations, max |rho err| {err:.2e}", flush=True)
    assert err <= 1e-6        # float32 rounding of values in [0.001, 1] is below 6e-8
    # element order: e = ex*nely + ey, the load sits at the right edge middle
    rho2 = rho.reshape(20, 10)
    np.testing.assert_array_equal(rho2[3], rho[30:40])


# ---------------------------------------------------------------------------
# the reference's loop body, imported under the reference's names
# ---------------------------------------------------------------------------


def test_reference_import_name_drives_loop():
    """optimizer/simp.py:54-61 imports, verbatim module names; the loop body of
    simp.py:144-197 through them, against the oracle. Between iterations the
    load vector is changed and restored (the solver recomputes |f| per call)."""
    import math

    from toptimization.fem import fields as F
    from toptimization.fem import kernels as K
    from toptimization.fem import solver as pcg
    from toptimization.material import compute_Ke
    from toptimization.mesh import build_edof, element_centers
    from toptimization.optimizer import filter as filt
    from toptimization.optimizer import oc as oc_mod
    prob = cantilever(40, 16, max_iterations=8, tolerance=0.0, cg_tolerance=1e-8, max_cg_iterations=20000)
    Ke = compute_Ke(prob.E, prob.nu, 2)
    F.allocate(prob.n_dofs, prob.n_elem, 8, (2 * math.ceil(prob.rmin) + 1) ** 2, Ke, build_edof(prob),
               prob.fixed_dofs, prob.force_dofs, prob.force_values)
    filt.precompute_filter(prob, element_centers(prob))
    F.rho.fill_(prob.volfrac)
    K.fill_scalar(F.u_prev, 0.0)
    cs = []
    for it in range(1, prob.max_iter + 1):
        filt.apply_density_filter(F.rho, F.rho_filt, F.filt_neighbors, F.filt_weights, F.filt_n_nb, prob.n_elem)
        K.copy_field(F.rho, F.rho_new)
        K.copy_field(F.rho_filt, F.rho)
        pcg.solve(prob.E_min, prob.penalty, 2, prob.max_cg_iter, prob.cg_tol, "jacobi", warm_start=it > 1)
        cs.append(K.compute_sensitivity(F.u, F.rho, F.dc, F.dv, F.Ke, F.edof, prob.E_min, prob.penalty,
                                        prob.n_elem, 2))
        K.copy_field(F.rho_new, F.rho)
        filt.apply_dc_filter_density_mode(prob.n_elem)
        oc_mod.oc_update(prob)
        if it == 4:
            rho_saved = F.rho.clone()
            # a zero load: 0 iterations and u = 0 (solver.py:58-64) ...
            f_saved = F.f.clone()
            F.f.zero_()
            u_prev = F.u_prev.clone()
            assert pcg.solve(prob.E_min, prob.penalty, 2, prob.max_cg_iter, prob.cg_tol, "jacobi") == 0
            assert float(F.u.abs().max()) == 0.0
            # ... a doubled load doubles u at the same density ...
            F.f.copy_(2.0 * f_saved)
            F.u_prev.copy_(u_prev)
            K.copy_field(F.rho_filt, F.rho)
            pcg.solve(prob.E_min, prob.penalty, 2, prob.max_cg_iter, prob.cg_tol, "jacobi", warm_start=False)
            u2 = F.u.clone()
            F.f.copy_(f_saved)
            pcg.solve(prob.E_min, prob.penalty, 2, prob.max_cg_iter, prob.cg_tol, "jacobi", warm_start=False)
            assert float((u2 - 2.0 * F.u).abs().max()) <= 1e-6 * float(u2.abs().max())
            # ... then the loop goes on with the original load and state
            F.rho.copy_(rho_saved)
            F.u_prev.copy_(u_prev)
    args, kw = oracle_args(prob)
    ref = O.simp_run(*args, **kw)
    rel = _rel(cs, ref["c"])
    print(f"toptimization.* loop: max rel c err {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-8
    np.testing.assert_allclose(F.rho.cpu().numpy(), ref["rho"], atol=1e-8)
