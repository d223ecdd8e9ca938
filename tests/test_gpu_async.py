"""Asynchronous SIMP driver (topt_plan_advance): every problem of a plan runs
its own OC iterations and PCG solves; results must be bitwise those of the
lockstep path (topt_plan_iteration) run on each problem alone, whatever the
chunk budget per call, the order of calls or the plan rebuilds between them."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solo(prob, precision):
    """Lockstep reference: one plan per problem, host stopping rule of
    optimizer/simp.py:217."""
    from toptimization_amd.engine import DesignState, Plan
    st = DesignState(prob, precision=precision)
    plan = Plan([st])
    it, n_cg = 0, 0
    while True:
        s = plan.iteration(sensitivity_filter=prob.filter_type == "sensitivity")[0]
        it += 1
        n_cg += s.n_cg
        if it >= prob.max_iter or (it > 5 and s.change < prob.tol):
            break
    plan.close()
    return it, n_cg, s, st.rho.cpu().numpy()


def _stream(n, seed, precision):
    from toptimization_amd.serving import synthetic_stream
    return synthetic_stream(n=n, rate=100.0, seed=seed, nelx_range=(32, 112), max_iter=14, tol=0.01,
                            cg_tolerance=1e-6, max_cg_iterations=20000,
                            preconditioner="block_jacobi_2x2", precision=precision)


@pytest.mark.parametrize("precision,max_chunks", [("mixed", 3), ("fp64", 1), ("mixed", 1 << 20)])
def test_advance_equals_solo_lockstep(precision, max_chunks):
    from toptimization_amd.engine import DesignState, Plan
    reqs = _stream(5, 11, precision)
    active = [{"rid": r.rid, "st": DesignState(r.problem, precision=precision), "phase": 0, "it": 0,
               "cg": 0, "last": None} for r in reqs]
    finished = {}
    calls = 0
    while active:
        print(f"advance call {calls}: {len(active)} active", flush=True)
        plan = Plan([a["st"] for a in active])
        phase = [a["phase"] for a in active]
        out, done_now, cg_now = plan.advance(phase, max_chunks=max_chunks)
        plan.close()
        calls += 1
        assert calls < 2000
        keep = []
        for a, ph, s, dn, cg in zip(active, phase, out, done_now, cg_now):
            assert ph in (0, 1, 3)
            a["phase"], a["it"], a["cg"] = ph, a["it"] + dn, a["cg"] + cg
            if dn:
                assert s.iter == a["it"]
                a["last"] = s
            if ph == 3:
                assert a["last"].done == 1
                finished[a["rid"]] = a
            else:
                keep.append(a)
        active = keep
    for r in reqs:
        it, n_cg, s, rho = _solo(r.problem, precision)
        a = finished[r.rid]
        assert a["it"] == it, r.rid
        assert a["cg"] == n_cg, r.rid
        got = a["last"]
        assert (got.compliance, got.volume, got.change, got.lam) == (s.compliance, s.volume, s.change, s.lam)
        assert (got.n_cg, got.cg_status, got.oc_steps, got.refines) == (s.n_cg, s.cg_status, s.oc_steps,
                                                                       s.refines)
        np.testing.assert_array_equal(a["st"].rho.cpu().numpy(), rho)


def test_advance_resumes_after_op_level_solve():
    """phase 2 in: a solve finished by topt_plan_solve gets its OC step first."""
    from toptimization_amd import _lib
    from toptimization_amd.engine import DesignState, Plan
    reqs = _stream(3, 5, "mixed")
    states = [DesignState(r.problem, precision="mixed") for r in reqs]
    plan = Plan(states)
    plan.filter(_lib.FILTER_DENSITY)
    n_it, status = plan.solve()
    assert all(st == _lib.ST_CONV for st in status)
    phase = [2] * len(states)
    its = [0] * len(states)
    while any(p != 3 for p in phase):
        out, done_now, _ = plan.advance(phase)
        its = [a + b for a, b in zip(its, done_now)]
    plan.close()
    for r, st, it in zip(reqs, states, its):
        ref_it, _, _, rho = _solo(r.problem, "mixed")
        assert it == ref_it
        np.testing.assert_array_equal(st.rho.cpu().numpy(), rho)


def test_advance_rejects_graph_plans_and_bad_phase():
    from toptimization_amd.engine import DesignState, Plan
    from toptimization_amd._lib import ToptError
    reqs = _stream(1, 2, "mixed")
    st = DesignState(reqs[0].problem, precision="mixed")
    plan = Plan([st], graph=True)
    with pytest.raises(ToptError):
        plan.advance([0])
    plan.close()
    plan = Plan([st])
    with pytest.raises(ToptError):
        plan.advance([7])
    plan.close()
