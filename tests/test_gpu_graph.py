"""HIP-graph replay of PCG chunks: every replay gives the eager result.

A plan with graph=True captures one 32-iteration chunk of slot launches and
replays it for every full chunk of every solve; between replays the program
does its own eager work (OC update, filters, status copies). Each OC
iteration's statistics and the final design must match the eager plan bit
for bit."""

import numpy as np
import pytest

from helpers import bridge, mbb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,preconditioner", [("mixed", "block_jacobi_2x2"), ("fp64", "jacobi"),
                                                      ("mixed", "block_jacobi")])
def test_graph_replay_matches_eager(precision, preconditioner):
    import torch
    from toptimization_amd.engine import DesignState, Plan
    probs = [mbb(48, 16, cg_tolerance=1e-8, max_cg_iterations=20000, preconditioner=preconditioner),
             bridge(40, 20, cg_tolerance=1e-8, max_cg_iterations=20000, preconditioner=preconditioner)]
    runs = []
    for graph in (False, True):
        states = [DesignState(p, precision=precision) for p in probs]
        plan = Plan(states, graph=graph)
        hist = []
        for _ in range(6):
            out = plan.iteration()
            torch.cuda.synchronize()
            hist.append([(o.compliance, o.n_cg, o.cg_status, o.change) for o in out])
        rho = [s.rho.cpu().numpy().copy() for s in states]
        plan.close()
        runs.append((hist, rho))
    (h0, r0), (h1, r1) = runs
    # the solves are long enough to use full chunks (more than 32 iterations)
    assert min(x[1] for it in h0 for x in it) > 64
    assert h0 == h1
    for a, b in zip(r0, r1):
        np.testing.assert_array_equal(a, b)
