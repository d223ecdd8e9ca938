"""Multigrid-preconditioned PCG (preconditioner "multigrid", no reference
counterpart): converged solutions against the oracle's PCG, far fewer
iterations than Jacobi, SIMP loops against the Jacobi path, and the
asynchronous driver bitwise equal to the lockstep one."""

import numpy as np
import pytest
from oracle import simp_np as O
from helpers import bridge, cantilever, mbb, rng_density

pytestmark = pytest.mark.gpu


def _solve(prob, precision, rho):
    from toptimization_amd.engine import DesignState, Plan
    st = DesignState(prob, precision=precision, rho0=rho)
    st.rhof.copy_(st.rho)
    plan = Plan([st])
    it, status = plan.solve(warm=0, use_filtered=True)
    plan.sensitivity(use_filtered=True)
    c = plan.stats()[0].compliance
    plan.close()
    return st.u.cpu().numpy(), it[0], status[0], c


# (56, 56): coarsest level 7x7 elements, 128 DOFs, half-bandwidth 19 (the
# largest the coarse factorization takes). The fused slot B + level-0 down
# sweep restricts on PCG tiles of 8 x 64 nodes, a tile to the coarse nodes
# over its own nodes: with odd nelx (nely) the coarse grid has one node
# column (row) past those of the last tile, which that tile also takes;
# (15, 31) and (63, 127) end the fine grid on the last node of a tile.
@pytest.mark.parametrize("shape", [(30, 12), (61, 23), (96, 32), (56, 56), (15, 31), (63, 127)])
@pytest.mark.parametrize("precision,tol,utol", [("fp64", 1e-10, 1e-7), ("mixed", 1e-8, 1e-5)])
def test_mg_pcg_matches_oracle(shape, precision, tol, utol):
    nelx, nely = shape
    print(f"multigrid PCG {precision} {nelx}x{nely}", flush=True)
    prob = mbb(nelx, nely, preconditioner="multigrid", cg_tolerance=tol, max_cg_iterations=20000)
    rho = rng_density(prob.n_elem, seed=nelx)
    u, it, status, c = _solve(prob, precision, rho)
    Ke, edof = O.ke_q4(prob.E, prob.nu), O.edof_q4(nelx, nely)
    fixed = O.fixed_mask(prob.n_dofs, prob.fixed_dofs)
    f = O.force_vector(prob.n_dofs, prob.force_dofs, prob.force_values, prob.fixed_dofs)
    u_ref, it_ref = O.pcg(f, O.scale(rho, prob.E_min, prob.penalty), Ke, edof, fixed,
                          max_iter=200000, tol=1e-12, preconditioner="jacobi")
    assert status == 1
    assert np.abs(u - u_ref).max() <= utol * np.abs(u_ref).max()
    assert c == pytest.approx(float(f @ u_ref), rel=10 * utol)
    assert it * 5 < it_ref


# The smoother's damping follows the element's Jacobi bound (omega = 0.97 *
# 2 / lambda_max(D_e^-1 Ke): 0.97 at nu = 0, 0.672 at 0.3, 0.569 at 0.45), so
# the V-cycle stays SPD for any Poisson ratio; a fixed 0.65 left the
# guaranteed range from nu = 0.35 on.
@pytest.mark.parametrize("nu", [0.0, 0.45])
@pytest.mark.parametrize("precision,tol,utol", [("fp64", 1e-10, 1e-7), ("mixed", 1e-8, 1e-5)])
def test_mg_pcg_other_poisson_ratios_match_oracle(nu, precision, tol, utol):
    nelx, nely = 96, 32
    print(f"multigrid PCG {precision} {nelx}x{nely} nu={nu}", flush=True)
    prob = mbb(nelx, nely, preconditioner="multigrid", cg_tolerance=tol, max_cg_iterations=20000, nu=nu)
    rho = rng_density(prob.n_elem, seed=7)
    u, it, status, c = _solve(prob, precision, rho)
    Ke, edof = O.ke_q4(prob.E, prob.nu), O.edof_q4(nelx, nely)
    fixed = O.fixed_mask(prob.n_dofs, prob.fixed_dofs)
    f = O.force_vector(prob.n_dofs, prob.force_dofs, prob.force_values, prob.fixed_dofs)
    u_ref, it_ref = O.pcg(f, O.scale(rho, prob.E_min, prob.penalty), Ke, edof, fixed,
                          max_iter=200000, tol=1e-12, preconditioner="jacobi")
    assert status == 1
    assert np.abs(u - u_ref).max() <= utol * np.abs(u_ref).max()
    assert c == pytest.approx(float(f @ u_ref), rel=10 * utol)
    assert it * 5 < it_ref


def test_mg_far_fewer_iterations_than_jacobi():
    prob_j = bridge(240, 80, preconditioner="jacobi", cg_tolerance=1e-6, max_cg_iterations=50000,
                    volume_fraction=0.3)
    rho = rng_density(prob_j.n_elem, seed=5, lo=0.001, hi=1.0)
    _, it_j, st_j, c_j = _solve(prob_j, "mixed", rho)
    prob_m = bridge(240, 80, preconditioner="multigrid", cg_tolerance=1e-6, max_cg_iterations=50000,
                    volume_fraction=0.3)
    _, it_m, st_m, c_m = _solve(prob_m, "mixed", rho)
    print(f"240x80 random density: Jacobi {it_j} CG, multigrid {it_m} CG", flush=True)
    assert st_j == st_m == 1
    assert it_m * 10 < it_j
    assert c_m == pytest.approx(c_j, rel=2e-6)


def test_mg_simp_loop_matches_jacobi_path():
    """60x20 cantilever, 30 OC iterations, FP64 PCG rtol 1e-8: multigrid and
    Jacobi preconditioning give the same designs (compliance within 1e-6)."""
    from toptimization_amd.engine import DesignState, Plan
    out = {}
    for pc in ("jacobi", "multigrid"):
        prob = cantilever(60, 20, preconditioner=pc, cg_tolerance=1e-8, max_cg_iterations=20000,
                          volume_fraction=0.5, filter_radius=1.5, max_iterations=30, tolerance=0.0)
        st = DesignState(prob, precision="fp64")
        plan = Plan([st])
        out[pc] = [plan.iteration()[0] for _ in range(30)]
        plan.close()
    cj = np.array([s.compliance for s in out["jacobi"]])
    cm = np.array([s.compliance for s in out["multigrid"]])
    assert np.abs(cm - cj).max() <= 1e-6 * np.abs(cj).max()
    assert sum(s.n_cg for s in out["multigrid"]) * 5 < sum(s.n_cg for s in out["jacobi"])


def test_mg_async_equals_lockstep():
    from toptimization_amd.engine import DesignState, Plan
    prob = mbb(96, 32, preconditioner="multigrid", cg_tolerance=1e-6, max_cg_iterations=20000,
               volume_fraction=0.5, filter_radius=2.0, max_iterations=12, tolerance=0.0)
    ref = DesignState(prob, precision="mixed")
    plan = Plan([ref])
    s_ref = [plan.iteration()[0] for _ in range(12)]
    plan.close()
    st = DesignState(prob, precision="mixed")
    phase = [0]
    its, cg = 0, 0
    while phase[0] != 3:
        plan = Plan([st])
        out, done, ncg = plan.advance(phase, max_chunks=5)
        plan.close()
        its += done[0]
        cg += ncg[0]
    assert its == 12
    assert cg == sum(s.n_cg for s in s_ref)
    assert out[0].compliance == s_ref[-1].compliance
    np.testing.assert_array_equal(st.rho.cpu().numpy(), ref.rho.cpu().numpy())


def test_mg_batch_equals_solo():
    """Ragged batch of multigrid problems (different depths and tail levels):
    every problem's iterates are bitwise those of its solo plan."""
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import synthetic_stream
    reqs = synthetic_stream(n=5, rate=100.0, seed=21, nelx_range=(32, 400), max_iter=6, tol=0.0,
                            cg_tolerance=1e-6, max_cg_iterations=20000, preconditioner="multigrid",
                            precision="mixed")
    states = [DesignState(r.problem, precision="mixed") for r in reqs]
    plan = Plan(states)
    out = [plan.iteration() for _ in range(6)]
    plan.close()
    for i, r in enumerate(reqs):
        print(f"multigrid batch member {i}: {r.problem.nelx}x{r.problem.nely}", flush=True)
        st = DesignState(r.problem, precision="mixed")
        p1 = Plan([st])
        solo = [p1.iteration()[0] for _ in range(6)]
        p1.close()
        assert [o[i].compliance for o in out] == [x.compliance for x in solo]
        assert [o[i].n_cg for o in out] == [x.n_cg for x in solo]
        np.testing.assert_array_equal(states[i].rho.cpu().numpy(), st.rho.cpu().numpy())


def test_mg_serving_async_equals_solo():
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import DesignServer, synthetic_stream
    reqs = synthetic_stream(n=6, rate=50.0, seed=4, nelx_range=(32, 160), max_iter=10, tol=0.01,
                            cg_tolerance=1e-6, max_cg_iterations=20000, preconditioner="multigrid",
                            precision="mixed")
    server = DesignServer(max_active=3, hbm_budget=1 << 30, precision="mixed", keep_rho=True)
    stats = server.run(reqs, clock="wall")
    assert stats.summary()["designs"] == 6
    for res in stats.results:
        st = DesignState(reqs[res.rid].problem, precision="mixed")
        plan = Plan([st])
        for _ in range(res.iterations):
            o = plan.iteration()[0]
        plan.close()
        assert o.compliance == res.compliance, res.rid
        np.testing.assert_array_equal(st.rho.cpu().numpy(), res.rho)


def test_mg_with_elemblock_in_one_plan():
    """Multigrid and element-block Jacobi share the three-slot iteration and
    can share a plan; each problem matches its solo run."""
    from toptimization_amd.engine import DesignState, Plan
    probs = [mbb(48, 16, preconditioner="multigrid", cg_tolerance=1e-8, max_cg_iterations=20000),
             mbb(40, 20, preconditioner="block_jacobi", cg_tolerance=1e-8, max_cg_iterations=20000)]
    states = [DesignState(p, precision="fp64") for p in probs]
    plan = Plan(states)
    both = [plan.iteration() for _ in range(3)]
    plan.close()
    for i, p in enumerate(probs):
        st = DesignState(p, precision="fp64")
        p1 = Plan([st])
        solo = [p1.iteration()[0] for _ in range(3)]
        p1.close()
        assert [b[i].compliance for b in both] == [x.compliance for x in solo]


def test_mg_deterministic_beyond_one_wave_of_workgroups():
    """1024x320 (390 PCG tiles, more workgroups than the GPU runs at once):
    the same design alone and next to another one in a batch (so its
    workgroups run at other times) gives bitwise the same iterates; a kernel
    that read values other workgroups of its own launch write would not."""
    from toptimization_amd.engine import DesignState, Plan
    big = bridge(1024, 320, preconditioner="multigrid", cg_tolerance=1e-6, max_cg_iterations=20000,
                 volume_fraction=0.3, filter_radius=5.0, max_iterations=3, tolerance=0.0, precision="mixed")
    other = mbb(256, 96, preconditioner="multigrid", cg_tolerance=1e-6, max_cg_iterations=20000,
                volume_fraction=0.5, filter_radius=3.0, max_iterations=3, tolerance=0.0, precision="mixed")
    a = DesignState(big, precision="mixed")
    p = Plan([a])
    solo = [p.iteration()[0] for _ in range(2)]
    p.close()
    b, c = DesignState(other, precision="mixed"), DesignState(big, precision="mixed")
    p = Plan([b, c])
    both = [p.iteration()[1] for _ in range(2)]
    p.close()
    print(f"1024x320 multigrid CG per OC iteration: {[s.n_cg for s in solo]}", flush=True)
    assert [s.n_cg for s in solo] == [s.n_cg for s in both]
    assert [s.compliance for s in solo] == [s.compliance for s in both]
    assert solo[0].n_cg <= 40


def test_mg_fp32_coarse_solve_near_null_support():
    """A 240x120 MBB of the serving stream whose y-support is one corner node:
    once the elements near it go void, the FP32 Galerkin chain turns the
    coarsest operator's near-null mode into pivots of either sign around 1e-3
    of the diagonal (a CPU check found an eigenvalue of -3.8e-6 where FP64
    has +3.2e-12). The modified pivots keep the coarse solve symmetric
    positive definite, so every solve converges in a few dozen iterations."""
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import synthetic_stream
    reqs = synthetic_stream(n=128, rate=20.0, seed=0, nelx_range=(64, 768), max_iter=100, tol=0.01,
                            cg_tolerance=1e-6, max_cg_iterations=20000, preconditioner="multigrid",
                            precision="mixed")
    prob = reqs[63].problem
    assert (prob.nelx, prob.nely) == (240, 120)
    st = DesignState(prob, precision="mixed")
    plan = Plan([st])
    out = [plan.iteration()[0] for _ in range(12)]
    plan.close()
    print(f"240x120 MBB multigrid CG per OC iteration: {[s.n_cg for s in out]}", flush=True)
    assert all(s.cg_status == 1 for s in out)
    assert max(s.n_cg for s in out) <= 200
    assert np.isfinite([s.compliance for s in out]).all()


@pytest.mark.parametrize("shape", [(10, 4), (7, 3), (2, 1)])
@pytest.mark.parametrize("precision,tol,utol", [("fp64", 1e-10, 1e-7), ("mixed", 1e-8, 1e-5)])
def test_mg_two_level_grids_match_oracle(shape, precision, tol, utol):
    """Grids whose level 1 is already the coarsest (<= 128 DOFs): the fused
    level-0 sweep writes the coarse right-hand side that the tail solves
    densely, with no level in between."""
    nelx, nely = shape
    print(f"multigrid PCG {precision} {nelx}x{nely} (two levels)", flush=True)
    prob = mbb(nelx, nely, preconditioner="multigrid", cg_tolerance=tol, max_cg_iterations=20000)
    rho = rng_density(prob.n_elem, seed=nelx + nely)
    u, it, status, c = _solve(prob, precision, rho)
    Ke, edof = O.ke_q4(prob.E, prob.nu), O.edof_q4(nelx, nely)
    fixed = O.fixed_mask(prob.n_dofs, prob.fixed_dofs)
    f = O.force_vector(prob.n_dofs, prob.force_dofs, prob.force_values, prob.fixed_dofs)
    u_ref, _ = O.pcg(f, O.scale(rho, prob.E_min, prob.penalty), Ke, edof, fixed,
                     max_iter=200000, tol=1e-12, preconditioner="jacobi")
    assert status == 1
    assert np.abs(u - u_ref).max() <= utol * np.abs(u_ref).max()
    assert c == pytest.approx(float(f @ u_ref), rel=10 * utol)
