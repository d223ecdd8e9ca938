"""The two-level down sweep (k_mg_down2: the last two down sweeps above the
multigrid tail of a one-problem plan in one launch) gives bitwise the results
of the two separate sweeps (k_mg_down_restrict twice): CG counts, compliance,
lambda, rho and u, on grids whose pair of levels is 1-2 (300x100), 2-3
(600x200, 640x213) and 3-4 (1200x400), mixed and FP64."""

import os

import numpy as np
import pytest

from helpers import bridge, mbb

pytestmark = pytest.mark.gpu


def _run(prob, n_it, separate):
    from toptimization_amd.engine import DesignState, Plan
    st = DesignState(prob)
    if separate:
        os.environ["TOPT_MG_NODOWN2"] = "1"
    try:
        plan = Plan([st])
    finally:
        os.environ.pop("TOPT_MG_NODOWN2", None)
    try:
        hist = [plan.iteration()[0] for _ in range(n_it)]
        return hist, st.rho.cpu().numpy(), st.u.cpu().numpy()
    finally:
        plan.close()


@pytest.mark.parametrize("case", ["mbb300x100-mixed", "bridge600x200-mixed", "mbb640x213-fp64",
                                  "bridge1200x400-mixed", "bridge1200x400-fp64"])
def test_down2_equals_separate_sweeps(case):
    name, prec = case.split("-")
    kw = dict(penalty=3.0, max_iterations=50, tolerance=0.0, cg_tolerance=1e-6, max_cg_iterations=20000,
              preconditioner="multigrid", precision=prec)
    if name == "mbb300x100":
        prob = mbb(300, 100, volume_fraction=0.5, filter_radius=4.0, **kw)
    elif name == "bridge600x200":
        prob = bridge(600, 200, volume_fraction=0.3, filter_radius=3.0, **kw)
    elif name == "mbb640x213":
        prob = mbb(640, 213, volume_fraction=0.5, filter_radius=4.0, **kw)
    else:
        prob = bridge(1200, 400, volume_fraction=0.3, filter_radius=6.0, **kw)
    n_it = 3
    print(f"[down2] {case}, {n_it} OC iterations", flush=True)
    ha, ra, ua = _run(prob, n_it, separate=False)
    hb, rb, ub = _run(prob, n_it, separate=True)
    for k, (a, b) in enumerate(zip(ha, hb)):
        assert a.n_cg == b.n_cg, (k, a.n_cg, b.n_cg)
        assert a.compliance == b.compliance, k
        assert a.lam == b.lam, k
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(ua, ub)
