"""The one-launch OC update (k_oc_coop: the whole bisection of
optimizer/oc.py:52-93 and the commit in one grid-barrier kernel) gives
bitwise the results of the multi-launch sequence (reset, chain pass, ten
subtree passes, commit): the same lambda, step count, volume, change and
rho, for one-problem plans whose fold takes one block (60x20), one group
(300x100: 15 blocks) and four groups (1200x400: 235 blocks), and for a
ragged batch of 265 logical blocks, more than the launch's 256 blocks, so
some blocks hold two and the loop ends on the barrier's run count."""

import os

import numpy as np
import pytest

from helpers import bridge, cantilever, mbb

pytestmark = pytest.mark.gpu


def _run(probs, n_it, multi):
    from toptimization_amd.engine import DesignState, Plan
    states = [DesignState(p) for p in probs]
    # the plan reads the switch once, when it is created
    if multi:
        os.environ["TOPT_OC_MULTI"] = "1"
    try:
        plan = Plan(states)
    finally:
        os.environ.pop("TOPT_OC_MULTI", None)
    try:
        hist = [plan.iteration() for _ in range(n_it)]
        rho = [s.rho.cpu().numpy() for s in states]
    finally:
        plan.close()
    return hist, rho


def _same(a, b):
    (ha, ra), (hb, rb) = a, b
    for k, (xa, xb) in enumerate(zip(ha, hb)):
        for i, (sa, sb) in enumerate(zip(xa, xb)):
            assert sa.oc_steps == sb.oc_steps and sa.oc_steps > 0, (k, i)
            assert sa.lam == sb.lam, (k, i)
            assert sa.volume == sb.volume, (k, i)
            assert sa.change == sb.change, (k, i)
            assert sa.compliance == sb.compliance, (k, i)
            assert sa.n_cg == sb.n_cg, (k, i)
            assert sa.iter == sb.iter == k + 1, (k, i)
    for i, (x, y) in enumerate(zip(ra, rb)):
        np.testing.assert_array_equal(x, y, err_msg=f"problem {i}")


@pytest.mark.parametrize("case", ["cantilever60x20", "mbb300x100", "bridge1200x400"])
def test_oc_coop_equals_multi_launch(case):
    if case == "cantilever60x20":
        prob = cantilever(60, 20, volume_fraction=0.5, filter_radius=1.5, penalty=3.0, max_iterations=50,
                          tolerance=0.0, cg_tolerance=1e-8, max_cg_iterations=5000, preconditioner="jacobi",
                          precision="fp64")
        n_it = 8
    elif case == "mbb300x100":
        prob = mbb(300, 100, volume_fraction=0.5, filter_radius=4.0, penalty=3.0, max_iterations=50,
                   tolerance=0.0, cg_tolerance=1e-6, max_cg_iterations=20000, preconditioner="multigrid",
                   precision="mixed")
        n_it = 6
    else:
        prob = bridge(1200, 400, volume_fraction=0.3, filter_radius=6.0, penalty=3.0, max_iterations=50,
                      tolerance=0.0, cg_tolerance=1e-6, max_cg_iterations=20000, preconditioner="multigrid",
                      precision="mixed")
        n_it = 3
    print(f"[oc_coop] {case}, {n_it} OC iterations", flush=True)
    _same(_run([prob], n_it, multi=False), _run([prob], n_it, multi=True))


def test_oc_coop_batch_equals_multi_launch():
    kw = dict(penalty=3.0, max_iterations=50, tolerance=0.0, cg_tolerance=1e-6, max_cg_iterations=20000,
              preconditioner="multigrid", precision="mixed")
    probs = [bridge(1200, 400, volume_fraction=0.3, filter_radius=6.0, **kw),
             mbb(300, 100, volume_fraction=0.5, filter_radius=4.0, **kw),
             cantilever(300, 100, volume_fraction=0.4, filter_radius=3.0, **kw)]
    print("[oc_coop] batch of 3 (265 OC blocks), 3 OC iterations", flush=True)
    _same(_run(probs, 3, multi=False), _run(probs, 3, multi=True))


def test_oc_coop_abort_path_reports_error():
    """A wait limit of one tick of the 100 MHz counter (set from the host)
    makes the blocks that wait for another block's bracket time out: the
    kernel raises the plan's abort word, every block leaves its loop, the
    commit is skipped (rho unchanged), and the iteration returns an error
    instead of hanging. The problem's reduction counters may be left mid-way,
    so neither the plan nor the state is used again."""
    import torch
    from toptimization_amd._lib import ToptError
    from toptimization_amd.engine import DesignState, Plan
    prob = mbb(600, 200, volume_fraction=0.5, filter_radius=4.0, penalty=3.0, max_iterations=50,
               tolerance=0.0, cg_tolerance=1e-6, max_cg_iterations=20000, preconditioner="multigrid",
               precision="mixed")
    st = DesignState(prob)
    os.environ["TOPT_OC_SPIN_TICKS"] = "1"
    try:
        plan = Plan([st])
    finally:
        os.environ.pop("TOPT_OC_SPIN_TICKS", None)
    rho0 = st.rho.clone()
    print("[oc_coop] 600x200 MBB (59 OC blocks), wait limit 1 tick", flush=True)
    try:
        with pytest.raises(ToptError, match="grid barrier timed out"):
            plan.iteration()
        assert torch.equal(st.rho, rho0)
    finally:
        plan.close()
