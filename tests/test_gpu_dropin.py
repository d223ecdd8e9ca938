"""The reference's module-level API (fields / kernels / fused_kernels /
preconditioner / solver / filter / oc), driven the way the reference's
optimizer/simp.py:99-197 drives it, against the FP64 oracle."""

import numpy as np
import pytest
import torch

from oracle import simp_np as O

from helpers import cantilever, mbb, oracle_args, rng_density

pytestmark = pytest.mark.gpu


def _allocate(prob):
    import math
    from toptimization_amd.fem import fields as F
    from toptimization_amd.material import compute_Ke
    from toptimization_amd.mesh import build_edof
    Ke = compute_Ke(prob.E, prob.nu, 2)
    max_nb = (2 * math.ceil(prob.rmin) + 1) ** 2
    F.allocate(prob.n_dofs, prob.n_elem, 8, max_nb, Ke, build_edof(prob), prob.fixed_dofs,
               prob.force_dofs, prob.force_values)
    return F


def test_reference_loop_through_dropin_api():
    from toptimization_amd.fem import kernels as K, solver as pcg
    from toptimization_amd.mesh import element_centers
    from toptimization_amd.optimizer import filter as filt, oc as oc_mod
    prob = cantilever(40, 16, max_iterations=8, tolerance=0.0, cg_tolerance=1e-8,
                      max_cg_iterations=20000)
    F = _allocate(prob)
    actual_max = filt.precompute_filter(prob, element_centers(prob))
    nb0, wt0, c0 = O.filter_tables_bruteforce(prob.nelx, prob.nely, prob.rmin)
    assert actual_max == c0.max()
    np.testing.assert_array_equal(F.filt_neighbors.cpu().numpy(), nb0)
    F.rho.fill_(prob.volfrac)
    K.fill_scalar(F.u_prev, 0.0)
    hist = {"c": [], "vol": [], "change": [], "n_cg": []}
    for it in range(1, prob.max_iter + 1):
        filt.apply_density_filter(F.rho, F.rho_filt, F.filt_neighbors, F.filt_weights, F.filt_n_nb,
                                  prob.n_elem)
        K.copy_field(F.rho, F.rho_new)
        K.copy_field(F.rho_filt, F.rho)
        n_cg = pcg.solve(prob.E_min, prob.penalty, 2, prob.max_cg_iter, prob.cg_tol, "jacobi",
                         warm_start=it > 1)
        c = K.compute_sensitivity(F.u, F.rho, F.dc, F.dv, F.Ke, F.edof, prob.E_min, prob.penalty,
                                  prob.n_elem, 2)
        K.copy_field(F.rho_new, F.rho)
        filt.apply_dc_filter_density_mode(prob.n_elem)
        change = oc_mod.oc_update(prob)
        vol = K.compute_volume(F.rho, prob.n_elem)
        for k, v in (("c", c), ("vol", vol), ("change", change), ("n_cg", n_cg)):
            hist[k].append(v)
    args, kw = oracle_args(prob)
    ref = O.simp_run(*args, **kw)
    np.testing.assert_allclose(hist["c"], ref["c"], rtol=1e-8)
    np.testing.assert_allclose(hist["vol"], ref["vol"], rtol=1e-10)
    assert np.all(np.abs(np.array(hist["n_cg"]) - ref["n_cg"]) <= 2)
    np.testing.assert_allclose(F.rho.cpu().numpy(), ref["rho"], atol=1e-8)


def test_dropin_pcg_pieces_match_oracle():
    from toptimization_amd.fem import fused_kernels as FK, kernels as K, preconditioner as PC
    prob = mbb(24, 10)
    F = _allocate(prob)
    Ke = O.ke_q4()
    edof = O.edof_q4(prob.nelx, prob.nely)
    fixed = O.fixed_mask(prob.n_dofs, prob.fixed_dofs)
    rho = rng_density(prob.n_elem, 3)
    F.rho.copy_(torch.as_tensor(rho))
    s = O.scale(rho, prob.E_min, prob.penalty)
    p = np.random.default_rng(4).normal(size=prob.n_dofs)
    F.p.copy_(torch.as_tensor(p))
    pAp = FK.matvec_and_dot(F.p, F.Ap, F.rho, F.Ke, F.edof, F.is_fixed, prob.E_min, prob.penalty,
                            prob.n_elem, 2)
    ref = O.matvec(p, s, Ke, edof, fixed)
    np.testing.assert_allclose(F.Ap.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)
    assert pAp == pytest.approx(p @ ref, rel=1e-11)
    K.compute_diagonal(F.diag, F.rho, F.Ke, F.edof, F.is_fixed, prob.E_min, prob.penalty, prob.n_elem, 2)
    d = O.diagonal(s, Ke, edof, fixed)
    np.testing.assert_allclose(F.diag.cpu().numpy(), d, rtol=1e-13)
    r = np.random.default_rng(5).normal(size=prob.n_dofs)
    r[fixed] = 0.0
    u = np.random.default_rng(6).normal(size=prob.n_dofs)
    F.r.copy_(torch.as_tensor(r)); F.u.copy_(torch.as_tensor(u))
    rz = FK.update_and_precond_jacobi(F.u, F.r, F.z, F.p, F.Ap, F.diag, F._dot_result, 0.25)
    r2 = r - 0.25 * ref
    np.testing.assert_allclose(F.u.cpu().numpy(), u + 0.25 * p, rtol=1e-14)
    np.testing.assert_allclose(F.z.cpu().numpy(), r2 / d, rtol=1e-12)
    assert rz == pytest.approx(r2 @ (r2 / d), rel=1e-12)
    assert F._dot_result[0].item() == pytest.approx(r2 @ r2, rel=1e-12)
    FK.update_p(F.z, F.p, 0.5)
    np.testing.assert_allclose(F.p.cpu().numpy(), r2 / d + 0.5 * p, rtol=1e-12)
    PC.setup(F.edof, F.diag_count, prob.n_elem, 2)
    np.testing.assert_array_equal(F.diag_count.cpu().numpy(), O.dof_count(edof, prob.n_dofs))
    F.r.copy_(torch.as_tensor(r))
    rz = PC.apply_and_rz(F.r, F.z, F.rho, F.Ke_inv, F.edof, F.diag_count, F.is_fixed, F._dot_result,
                         prob.E_min, prob.penalty, F.KE_ALPHA_REG, prob.n_elem, 2)
    Ki, a = O.ke_inverse_regularised(Ke)
    zref = O.elemblock_apply(r, s, Ki, a, edof, O.dof_count(edof, prob.n_dofs), fixed)
    np.testing.assert_allclose(F.z.cpu().numpy(), zref, rtol=1e-11, atol=1e-12 * np.abs(zref).max())
    assert rz == pytest.approx(r @ zref, rel=1e-11)
    assert K.dot_product(F.r, F.z) == pytest.approx(r @ zref, rel=1e-11)
    assert K.max_density_change(F.rho, F.rho_filt, prob.n_elem) == pytest.approx(np.abs(rho).max())
