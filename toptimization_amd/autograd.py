"""Differentiable compliance c(rho) on the GPU engine.

``ComplianceFunction.apply(rho, model)``: forward = density filter -> PCG ->
compliance (the reference loop body, ``optimizer/simp.py:144-168``);
backward = the element sensitivity kernel (``fem/kernels.py:234-272``) on
the forward's displacement (compliance is self-adjoint, so no adjoint solve)
followed by the filter's chain rule on the GPU:

* ``chain="exact"``     H^T (dc / (Hs + 1e-20)), the true gradient;
* ``chain="reference"`` H dc / (Hs + 1e-20), the reference's projection
  (``optimizer/filter.py:127-149``), equal to the exact one away from the
  boundary band of width rmin.

``TopologyModel`` wraps one problem's device state so that
``model.compliance(rho)`` can be used inside any torch optimisation loop.
"""

from __future__ import annotations

import torch

from . import _lib, ops
from .engine import DesignState, Plan
from .problem import Problem


class TopologyModel:
    """One design problem resident on the GPU, differentiable in rho."""

    def __init__(self, problem: Problem, precision: str | None = None,
                 preconditioner: str | None = None, chain: str = "exact"):
        if chain not in ("exact", "reference"):
            raise ValueError("chain must be 'exact' or 'reference'")
        if problem.filter_type != "density":
            raise NotImplementedError("the differentiable model uses the density filter")
        self.problem = problem
        self.chain = chain
        self.state = DesignState(problem, precision=precision, preconditioner=preconditioner)
        self.plan = Plan([self.state])
        self.last_n_cg = 0

    def compliance(self, rho: torch.Tensor) -> torch.Tensor:
        return ComplianceFunction.apply(rho, self)

    def forward_raw(self, rho: torch.Tensor, warm: bool = True,
                    energy: bool = False) -> tuple[float, torch.Tensor]:
        """Filter, solve and sensitivity for rho; returns (c, dc_phys copy).

        energy=True returns 2 f.u - u^T K u instead of c = u^T K u: equal at
        the exact solution, but its error is quadratic in the PCG error
        (finite-difference checks need that)."""
        st = self.state
        st.rho.copy_(rho.detach().to(torch.float64).reshape(-1))
        self.plan.filter(_lib.FILTER_DENSITY)
        it, status = self.plan.solve(warm=1 if warm else 0, use_filtered=True)
        if status[0] not in (_lib.ST_CONV, _lib.ST_ZERO):
            raise RuntimeError(f"PCG did not converge (status {status[0]}, {it[0]} iterations)")
        self.last_n_cg = it[0]
        self.plan.sensitivity(use_filtered=True)
        c = self.plan.stats()[0].compliance
        if energy:
            fu = float(ops.blas1("dot", st.u.numel(), x=st.f, y=st.u, reduce=True)[0].item())
            c = 2.0 * fu - c
        return c, st.dc.clone()

    def chain_rule(self, dc_phys: torch.Tensor) -> torch.Tensor:
        p = self.problem
        mode = _lib.FILTER_TRANSPOSE if self.chain == "exact" else _lib.FILTER_DCAVG
        return ops.filter_apply(p.nelx, p.nely, p.rmin, mode, dc=dc_phys.contiguous())

    def close(self) -> None:
        self.plan.close()


class ComplianceFunction(torch.autograd.Function):
    """c = f^T u with K(filter(rho)) u = f; d c / d rho by the sensitivity kernel."""

    @staticmethod
    def forward(ctx, rho: torch.Tensor, model: TopologyModel) -> torch.Tensor:
        c, dc_phys = model.forward_raw(rho)
        ctx.model = model
        ctx.shape, ctx.dtype = rho.shape, rho.dtype
        ctx.save_for_backward(dc_phys)
        return torch.tensor(c, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad_c: torch.Tensor):
        (dc_phys,) = ctx.saved_tensors
        g = ctx.model.chain_rule(dc_phys)
        # the gradient has rho's shape and dtype (flat, or (nelx, nely) as
        # simp.run returns it)
        return (grad_c * g).reshape(ctx.shape).to(ctx.dtype), None
