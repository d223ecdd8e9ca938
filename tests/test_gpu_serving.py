"""Continuous batching: requests that arrive while others run are admitted as
slots / memory free up; every design is bitwise the same as when it is solved
alone."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["async", "lockstep"])
def test_stream_results_equal_alone_runs(mode):
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import DesignServer, synthetic_stream
    reqs = synthetic_stream(n=7, rate=50.0, seed=3, nelx_range=(32, 96), max_iter=12, tol=0.01,
                            cg_tolerance=1e-6, max_cg_iterations=20000,
                            preconditioner="block_jacobi_2x2", precision="mixed")
    server = DesignServer(max_active=3, hbm_budget=1 << 30, precision="mixed", keep_rho=True, mode=mode)
    stats = server.run(reqs, clock="wall")
    s = stats.summary()
    assert s["designs"] == 7 and s["max_active"] == 3
    assert all(r.finished >= r.admitted >= r.arrival - 1e-9 for r in stats.results)
    # some request was admitted while others were still running
    overlap = any(b.admitted < a.finished and b.rid != a.rid and b.admitted > a.admitted
                  for a in stats.results for b in stats.results)
    assert overlap
    for res in stats.results:
        prob = reqs[res.rid].problem
        st = DesignState(prob, precision="mixed")
        plan = Plan([st])
        for _ in range(res.iterations):
            out = plan.iteration()[0]
        plan.close()
        assert out.compliance == res.compliance, res.rid
        assert (out.volume, out.change) == (res.volume, res.change), res.rid
        np.testing.assert_array_equal(st.rho.cpu().numpy(), res.rho)


def test_admission_respects_memory_budget():
    from toptimization_amd.serving import DesignServer, footprint_bytes, synthetic_stream
    reqs = synthetic_stream(n=4, rate=1000.0, seed=5, nelx_range=(48, 48), max_iter=3, tol=0.0,
                            cg_tolerance=1e-6, max_cg_iterations=20000, precision="fp64")
    one = footprint_bytes(reqs[0].problem, "fp64")
    server = DesignServer(max_active=8, hbm_budget=int(2.5 * one), precision="fp64")
    stats = server.run(reqs, clock="instant")
    assert stats.summary()["designs"] == 4
    assert stats.max_active == 2


def test_cli_serve_jsonl(tmp_path):
    """`toptimization serve requests.jsonl`: JSON request format in, one
    density file per design and a summary out."""
    import json

    from toptimization_amd.cli import main
    reqs = []
    for i, (nelx, nely) in enumerate([(24, 8), (32, 12)]):
        reqs.append({"id": 100 + i, "arrival": 0.0,
                     "problem": {"problem": {"name": f"c{i}"}, "domain": {"nelx": nelx, "nely": nely},
                                 "boundary_conditions": {"fixed": [
                                     {"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x", "y"]}]},
                                 "loads": [{"nodes": {"type": "point", "location": [nelx, 0]},
                                            "force": [0.0, -1.0]}],
                                 "optimization": {"volume_fraction": 0.5, "filter_radius": 1.5,
                                                  "max_iterations": 6, "tolerance": 0.0}}})
    f = tmp_path / "reqs.jsonl"
    f.write_text("\n".join(json.dumps(r) for r in reqs) + "\n")
    out = tmp_path / "out"
    assert main(["serve", str(f), "--output", str(out), "--max-active", "2"]) == 0
    summary = json.loads((out / "summary.json").read_text())
    assert summary["summary"]["designs"] == 2
    assert sorted(d["id"] for d in summary["designs"]) == [100, 101]
    for d in summary["designs"]:
        rho = np.load(out / f"{d['id']}_density.npy")
        assert rho.shape == (d["nelx"], d["nely"])
        assert d["iterations"] == 6
        assert abs(rho.mean() - 0.5) < 1e-3


@pytest.mark.parametrize("mode", ["async", "lockstep", "auto"])
def test_stream_mixing_plan_classes(mode):
    """Requests naming their own preconditioner (three-slot multigrid,
    two-slot Jacobi and 2x2 block Jacobi) and filter type in one stream: the
    server runs one plan per class, and every design equals its solo run."""
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd.serving import DesignServer, plan_class, synthetic_stream
    reqs = synthetic_stream(n=6, rate=100.0, seed=4, nelx_range=(32, 80), max_iter=8, tol=0.01,
                            cg_tolerance=1e-6, max_cg_iterations=20000, precision="mixed")
    pcs = ["multigrid", "jacobi", "block_jacobi_2x2", "multigrid", "jacobi", "multigrid"]
    for r, pc in zip(reqs, pcs):
        r.problem.preconditioner = pc
    reqs[3].problem.filter_type = "sensitivity"
    assert len({plan_class(r.problem) for r in reqs}) == 3
    server = DesignServer(max_active=4, hbm_budget=1 << 30, precision="mixed", keep_rho=True, mode=mode)
    stats = server.run(reqs, clock="instant")
    assert stats.summary()["designs"] == 6
    for res in stats.results:
        prob = reqs[res.rid].problem
        st = DesignState(prob, precision="mixed")
        plan = Plan([st])
        for _ in range(res.iterations):
            out = plan.iteration(sensitivity_filter=prob.filter_type == "sensitivity")[0]
        plan.close()
        assert (out.compliance, out.volume, out.change) == (res.compliance, res.volume, res.change), res.rid
        np.testing.assert_array_equal(st.rho.cpu().numpy(), res.rho)
