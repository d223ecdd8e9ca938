"""Pins for the 3-D (H8) part of the FP64 oracle against independent
mathematics: rigid-body modes, the constant-strain patch test, the
reference's own Ke tests, explicit connectivity, brute-force neighbour
tables, SciPy sparse assembly and spsolve, finite differences."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import simp_np as O
from toptimization_amd.material import compute_Ke_3d
from toptimization_amd.mesh import build_edof_3d
from toptimization_amd.problem import make_problem

from helpers import rng_density

CORNERS = np.array(O.H8_NODES, dtype=float)


def cantilever3(nelx=6, nely=3, nelz=4, **kw):
    return make_problem(nelx, nely, nelz=nelz,
                        fixed=[{"nodes": {"type": "face", "face": "left"}, "dofs": ["x", "y", "z"]}],
                        loads=[{"nodes": {"type": "point", "location": [nelx, 0, nelz // 2]},
                                "force": [0.0, -1.0, 0.0]}], name="cantilever3", **kw)


def assemble(s, Ke, edof, n):
    d = Ke.shape[0]
    rows = np.repeat(edof, d, axis=1).ravel()
    cols = np.tile(edof, (1, d)).ravel()
    vals = (s[:, None, None] * Ke[None]).ravel()
    return sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()


def bc_operator(K, fixed):
    K = K.tolil()
    for i in np.flatnonzero(fixed):
        K.rows[i] = [i]
        K.data[i] = [1.0]
    return K.tocsr()


def test_ke_h8_rigid_modes_and_reference_properties():
    Ke = O.ke_h8(1.0, 0.3)
    # the reference's tests (tests/test_material.py:35-52)
    assert Ke.shape == (24, 24)
    np.testing.assert_allclose(Ke, Ke.T, atol=1e-12)
    assert np.linalg.eigvalsh(Ke).min() >= -1e-8
    assert np.linalg.matrix_rank(Ke, tol=1e-8) == 18
    # three translations and three infinitesimal rotations are in the null space
    x = CORNERS
    modes = []
    for c in range(3):
        m = np.zeros((8, 3)); m[:, c] = 1.0; modes.append(m.ravel())
    for a, b in ((0, 1), (1, 2), (0, 2)):
        m = np.zeros((8, 3)); m[:, a] = -x[:, b]; m[:, b] = x[:, a]; modes.append(m.ravel())
    for m in modes:
        assert np.abs(Ke @ m).max() < 1e-13
    # the product's quadrature equals the oracle's
    assert np.abs(Ke - compute_Ke_3d(1.0, 0.3)).max() < 1e-14
    assert np.abs(O.ke_h8(2.0, 0.25) - compute_Ke_3d(2.0, 0.25)).max() < 1e-14


def test_ke_h8_constant_strain_patch_test():
    """u = G x on the unit cube: the trilinear element reproduces the linear
    field, so the element energy 1/2 u^T Ke u equals 1/2 eps : sigma over the
    unit volume, with eps the symmetric part of G (the skew part costs
    nothing)."""
    E, nu = 1.7, 0.28
    Ke = O.ke_h8(E, nu)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    rng = np.random.default_rng(7)
    for _ in range(5):
        G = rng.normal(size=(3, 3))
        u = (CORNERS @ G.T).ravel()
        eps = 0.5 * (G + G.T)
        energy = 0.5 * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
        assert 0.5 * u @ Ke @ u == pytest.approx(energy, rel=1e-13)


def test_edof_h8_explicit_and_product():
    # element 0 of a 2 x 3 x 4 grid: node (ix, iy, iz) -> ix*20 + iy*5 + iz
    e0 = O.edof_h8(2, 3, 4)[0]
    nodes = [0, 20, 25, 5, 1, 21, 26, 6]
    np.testing.assert_array_equal(e0, np.array([[3 * n, 3 * n + 1, 3 * n + 2] for n in nodes]).ravel())
    prob = cantilever3(2, 3, 4)
    np.testing.assert_array_equal(build_edof_3d(prob), O.edof_h8(2, 3, 4))
    # the last element touches the last node
    assert O.edof_h8(2, 3, 4).max() == 3 * (3 * 4 * 5) - 1


@pytest.mark.parametrize("rmin", [1.0, 1.5, 2.0, 2.6])
def test_filter_tables_3d_bruteforce(rmin):
    nb0, wt0, c0 = O.filter_tables_bruteforce_3d(5, 3, 4, rmin)
    nb1, wt1, c1 = O.filter_tables_3d(5, 3, 4, rmin)
    np.testing.assert_array_equal(nb0, nb1)
    np.testing.assert_array_equal(wt0.view(np.uint32), wt1.view(np.uint32))
    np.testing.assert_array_equal(c0, c1)
    # an element deep inside sees the whole ball: count of integer offsets with |d| < rmin
    r = int(np.ceil(rmin))
    ball = sum(1 for dx in range(-r, r + 1) for dy in range(-r, r + 1) for dz in range(-r, r + 1)
               if dx * dx + dy * dy + dz * dz < rmin * rmin)
    assert c0.max() <= ball


def _setup3(prob, seed=0):
    Ke = O.ke_h8(prob.E, prob.nu)
    edof = O.edof_h8(prob.nelx, prob.nely, prob.nelz)
    fixed = O.fixed_mask(prob.n_dofs, prob.fixed_dofs)
    f = O.force_vector(prob.n_dofs, prob.force_dofs, prob.force_values, prob.fixed_dofs)
    rho = rng_density(prob.n_elem, seed, 0.05, 1.0)
    return Ke, edof, fixed, f, rho, O.scale(rho, prob.E_min, prob.penalty)


def test_matvec_diag_elemblock_3d_against_assembly():
    prob = cantilever3(4, 3, 3)
    Ke, edof, fixed, f, rho, s = _setup3(prob, 1)
    K = assemble(s, Ke, edof, prob.n_dofs)
    A = bc_operator(K, fixed)
    p = np.random.default_rng(2).normal(size=prob.n_dofs)
    np.testing.assert_allclose(O.matvec(p, s, Ke, edof, fixed), A @ p, rtol=1e-12, atol=1e-12)
    d = K.diagonal().copy()
    d[fixed] = 1.0
    np.testing.assert_allclose(O.diagonal(s, Ke, edof, fixed), d, rtol=1e-13)
    # element-block Jacobi as a dense operator: symmetric and positive definite
    # on the free DOFs, zero on fixed rows
    Ki, a = O.ke_inverse_regularised(Ke)
    cnt = O.dof_count(edof, prob.n_dofs)
    M = np.stack([O.elemblock_apply(np.eye(prob.n_dofs)[:, j], s, Ki, a, edof, cnt, fixed)
                  for j in range(prob.n_dofs)], axis=1)
    Mf = M[np.ix_(~fixed, ~fixed)]
    np.testing.assert_allclose(Mf, Mf.T, atol=1e-12 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (Mf + Mf.T)).min() > 0
    assert np.abs(M[fixed]).max() == 0.0
    # the assembled form of the same operator, built element by element
    w = s / np.maximum(s, a) ** 2
    C = sp.diags(1.0 / np.sqrt(cnt))
    Mref = (C @ assemble(w, Ki, edof, prob.n_dofs) @ C).toarray()
    Mref[fixed] = 0.0
    np.testing.assert_allclose(M, Mref, rtol=1e-12, atol=1e-14 * np.abs(Mref).max())
    # interior node counts: 8 elements share an interior node
    ix = iy = iz = 1
    n = (ix * (prob.nely + 1) + iy) * (prob.nelz + 1) + iz
    assert cnt[3 * n] == 8


@pytest.mark.parametrize("pc", ["jacobi", "none", "block_jacobi"])
def test_pcg_3d_converges_to_spsolve(pc):
    prob = cantilever3(4, 3, 3)
    Ke, edof, fixed, f, rho, s = _setup3(prob, 3)
    A = bc_operator(assemble(s, Ke, edof, prob.n_dofs), fixed)
    u_ref = spla.spsolve(A.tocsc(), f)
    u, n = O.pcg(f, s, Ke, edof, fixed, max_iter=5000, tol=1e-12, preconditioner=pc)
    assert 0 < n < 5000
    np.testing.assert_allclose(u, u_ref, rtol=1e-8, atol=1e-10 * np.abs(u_ref).max())
    assert f @ u == pytest.approx(u @ (A @ u), rel=1e-10)


def test_sensitivity_3d_finite_differences():
    prob = cantilever3(3, 2, 2)
    Ke, edof, fixed, f, rho, _ = _setup3(prob, 4)

    def comp(r):
        s = O.scale(r, prob.E_min, prob.penalty)
        A = bc_operator(assemble(s, Ke, edof, prob.n_dofs), fixed)
        return float(f @ spla.spsolve(A.tocsc(), f))

    s = O.scale(rho, prob.E_min, prob.penalty)
    A = bc_operator(assemble(s, Ke, edof, prob.n_dofs), fixed)
    u = spla.spsolve(A.tocsc(), f)
    c, dc, dv = O.sensitivity(u, rho, Ke, edof, prob.E_min, prob.penalty)
    assert c == pytest.approx(f @ u, rel=1e-10)
    h = 1e-6
    for j in (0, 5, 11):
        rp, rm = rho.copy(), rho.copy()
        rp[j] += h; rm[j] -= h
        assert dc[j] == pytest.approx((comp(rp) - comp(rm)) / (2 * h), rel=1e-5)


def test_simp_3d_loop_behaves():
    prob = cantilever3(8, 4, 4, volume_fraction=0.3, filter_radius=1.5, max_iterations=12, tolerance=0.0,
                       cg_tolerance=1e-8, max_cg_iterations=20000)
    h = O.simp_run_3d(prob.nelx, prob.nely, prob.nelz, prob.fixed_dofs, prob.force_dofs, prob.force_values,
                      volfrac=0.3, rmin=1.5, max_iter=12, tol=0.0, cg_tol=1e-8, max_cg=20000)
    assert len(h["c"]) == 12
    assert h["c"][-1] < h["c"][0]
    assert abs(h["vol"][-1] - 0.3) < 5e-3
    assert h["rho"].shape == (prob.n_elem,)
