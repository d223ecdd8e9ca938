"""Serving request format (JSON lines) -> Problems, host side only."""

import json

import numpy as np
import pytest

from toptimization_amd.problem import make_problem
from toptimization_amd.serving import load_requests, request_from_json, synthetic_stream


def _req(i, nelx=40, nely=10, arrival=0.5):
    return {"id": i, "arrival": arrival,
            "problem": {"problem": {"name": f"r{i}", "dimension": 2},
                        "domain": {"nelx": nelx, "nely": nely},
                        "boundary_conditions": {"fixed": [
                            {"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x", "y"]}]},
                        "loads": [{"nodes": {"type": "point", "location": [nelx, nely // 2]},
                                   "force": [0.0, -1.0]}],
                        "optimization": {"volume_fraction": 0.4, "filter_radius": 1.5, "max_iterations": 7},
                        "solver": {"precision": "mixed", "preconditioner": "block_jacobi_2x2"}}}


def test_request_matches_programmatic_problem():
    r = request_from_json(json.dumps(_req(3)))
    ref = make_problem(40, 10, fixed=[{"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x", "y"]}],
                       loads=[{"nodes": {"type": "point", "location": [40, 5]}, "force": [0.0, -1.0]}],
                       name="r3", volume_fraction=0.4, filter_radius=1.5, max_iterations=7,
                       precision="mixed", preconditioner="block_jacobi_2x2")
    assert r.rid == 3 and r.arrival == 0.5
    p = r.problem
    np.testing.assert_array_equal(p.fixed_dofs, ref.fixed_dofs)
    np.testing.assert_array_equal(p.force_dofs, ref.force_dofs)
    np.testing.assert_array_equal(p.force_values, ref.force_values)
    assert (p.volfrac, p.rmin, p.max_iter, p.precision, p.preconditioner) == \
        (ref.volfrac, ref.rmin, ref.max_iter, ref.precision, ref.preconditioner)


def test_jsonl_file_and_defaults(tmp_path):
    f = tmp_path / "reqs.jsonl"
    lines = [json.dumps(_req(10)), "", json.dumps({"problem": _req(0)["problem"]})]
    f.write_text("\n".join(lines) + "\n")
    reqs = load_requests(f)
    assert [r.rid for r in reqs] == [10, 2]          # missing id -> line number
    assert reqs[1].arrival == 0.0


def test_invalid_requests_raise():
    with pytest.raises(ValueError):
        request_from_json({"id": 1})
    bad = _req(1)
    bad["problem"]["domain"]["nelx"] = -4
    with pytest.raises(ValueError):
        request_from_json(bad)


def test_synthetic_stream_is_seeded():
    a = synthetic_stream(n=5, seed=4)
    b = synthetic_stream(n=5, seed=4)
    assert [(r.problem.nelx, r.problem.nely, r.arrival) for r in a] == \
        [(r.problem.nelx, r.problem.nely, r.arrival) for r in b]
    assert all(r.problem.nelx % 16 == 0 and 64 <= r.problem.nelx <= 768 for r in a)


def test_plan_class_groups_by_slot_count_and_filter():
    from toptimization_amd.serving import plan_class
    kw = dict(fixed=[{"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x", "y"]}],
              loads=[{"nodes": {"type": "point", "location": [20, 5]}, "force": [0.0, -1.0]}])
    keys = {}
    for pc in ("none", "jacobi", "block_jacobi", "block_jacobi_2x2", "multigrid"):
        for ft in ("density", "sensitivity"):
            keys[pc, ft] = plan_class(make_problem(20, 10, preconditioner=pc, filter_type=ft, **kw))
    assert keys["jacobi", "density"] == keys["none", "density"] == keys["block_jacobi_2x2", "density"]
    assert keys["multigrid", "density"] == keys["block_jacobi", "density"] != keys["jacobi", "density"]
    assert keys["jacobi", "density"] != keys["jacobi", "sensitivity"]
    assert len(set(keys.values())) == 4
    # a server-wide preconditioner overrides the request's
    p = make_problem(20, 10, preconditioner="jacobi", **kw)
    assert plan_class(p, "multigrid") == keys["multigrid", "density"]


def test_server_mode_resolution():
    """mode='auto' advances multigrid designs in lockstep and the others
    asynchronously; the explicit modes apply to every design."""
    from toptimization_amd.serving import DesignServer
    kw = dict(fixed=[{"nodes": {"type": "edge", "edge": "left"}, "dofs": ["x", "y"]}],
              loads=[{"nodes": {"type": "point", "location": [20, 5]}, "force": [0.0, -1.0]}])
    mg = make_problem(20, 10, preconditioner="multigrid", **kw)
    jac = make_problem(20, 10, preconditioner="jacobi", **kw)
    auto = DesignServer()
    assert auto.mode == "auto"
    assert auto._lockstep(mg) and not auto._lockstep(jac)
    assert DesignServer(preconditioner="multigrid")._lockstep(jac)
    assert not DesignServer(mode="async")._lockstep(mg)
    assert DesignServer(mode="lockstep")._lockstep(jac)
    with pytest.raises(ValueError):
        DesignServer(mode="eager")
