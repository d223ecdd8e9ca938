"""FP64 NumPy restatement of the reference SIMP hot path (TEST INFRASTRUCTURE).

Only tests, the smoke check and the benchmark's CPU-baseline leg use this
module. Every function cites the reference file:line it restates
(paths relative to the reference ``src/toptimization/``). Arithmetic is FP64
throughout (the reference runs FP32, ``optimizer/simp.py:48``); data the
reference stores in FP32 (filter weights) is rounded to FP32 exactly as the
reference stores it and then used in FP64.

Pins (tests/test_oracle.py): closed-form Ke, reference mesh/problem tests,
SciPy sparse assembly + spsolve, dense filter matrices, brute-force tables,
finite differences, brute-force OC replay. Functions without a reference
counterpart (2x2 nodal block Jacobi, exact filter transpose) are pinned by
the same mathematical checks.
"""

from __future__ import annotations

import math

import numpy as np

# --------------------------------------------------------------------------
# setup: Ke, connectivity, BC vectors
# --------------------------------------------------------------------------


def ke_q4(E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """Closed-form Q4 plane-stress Ke (Sigmund 2001, "99-line" k-vector).

    The reference computes the same matrix by 2x2 Gauss quadrature
    (``material.py:20-86``); equality to 1e-15 is a test.
    """
    k = np.array([0.5 - nu / 6, 0.125 + nu / 8, -0.25 - nu / 12, -0.125 + 3 * nu / 8,
                  -0.25 + nu / 12, -0.125 - nu / 8, nu / 6, 0.125 - 3 * nu / 8])
    idx = np.array([[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2],
                    [2, 7, 0, 5, 6, 3, 4, 1], [3, 6, 5, 0, 7, 2, 1, 4],
                    [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6],
                    [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]])
    return E / (1.0 - nu * nu) * k[idx]


def edof_q4(nelx: int, nely: int) -> np.ndarray:
    """Element -> DOF map, restating the loop of ``mesh.py:33-65``."""
    out = np.zeros((nelx * nely, 8), dtype=np.int64)
    for ex in range(nelx):
        for ey in range(nely):
            n1 = ex * (nely + 1) + ey
            n2 = (ex + 1) * (nely + 1) + ey
            out[ex * nely + ey] = [2 * n1, 2 * n1 + 1, 2 * n2, 2 * n2 + 1,
                                   2 * n2 + 2, 2 * n2 + 3, 2 * n1 + 2, 2 * n1 + 3]
    return out


H8_NODES = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def ke_h8(E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """24x24 H8 stiffness of the unit cube by 2x2x2 Gauss quadrature
    (``material.py:89-171``): isotropic D (Lame lambda, mu), trilinear shape
    functions written out node by node, Jacobian 1/2 per axis. Pinned by the
    six rigid-body modes and the constant-strain patch test, not by
    comparison with the reference's code."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    for i in range(3):
        for j in range(3):
            D[i, j] = lam + (2.0 * mu if i == j else 0.0)
        D[3 + i, 3 + i] = mu
    g = 1.0 / math.sqrt(3.0)
    K = np.zeros((24, 24))
    for xi in (-g, g):
        for eta in (-g, g):
            for zeta in (-g, g):
                B = np.zeros((6, 24))
                for a, (cx, cy, cz) in enumerate(H8_NODES):
                    sx, sy, sz = 2 * cx - 1, 2 * cy - 1, 2 * cz - 1
                    # d/dx = 2 d/dxi on the unit cube
                    nx = sx * (1 + sy * eta) * (1 + sz * zeta) / 4.0
                    ny = sy * (1 + sx * xi) * (1 + sz * zeta) / 4.0
                    nz = sz * (1 + sx * xi) * (1 + sy * eta) / 4.0
                    c = 3 * a
                    B[0, c] = nx
                    B[1, c + 1] = ny
                    B[2, c + 2] = nz
                    B[3, c], B[3, c + 1] = ny, nx          # gamma_xy
                    B[4, c + 1], B[4, c + 2] = nz, ny      # gamma_yz
                    B[5, c], B[5, c + 2] = nz, nx          # gamma_xz
                K += B.T @ D @ B / 8.0
    return K


def edof_h8(nelx: int, nely: int, nelz: int) -> np.ndarray:
    """3-D element -> DOF map, restating the loop of ``mesh.py:68-112``."""
    out = np.zeros((nelx * nely * nelz, 24), dtype=np.int64)
    for ex in range(nelx):
        for ey in range(nely):
            for ez in range(nelz):
                e = ex * nely * nelz + ey * nelz + ez
                for k, (dx, dy, dz) in enumerate(H8_NODES):
                    n = (ex + dx) * (nely + 1) * (nelz + 1) + (ey + dy) * (nelz + 1) + (ez + dz)
                    out[e, 3 * k:3 * k + 3] = [3 * n, 3 * n + 1, 3 * n + 2]
    return out


def force_vector(n_dofs: int, force_dofs, force_values, fixed_dofs) -> np.ndarray:
    """``fields.py:151-155``: duplicate loads summed, fixed entries zeroed."""
    f = np.zeros(n_dofs)
    for d, v in zip(force_dofs, force_values):
        f[int(d)] += float(v)
    f[np.asarray(fixed_dofs, dtype=np.int64)] = 0.0
    return f


def fixed_mask(n_dofs: int, fixed_dofs) -> np.ndarray:
    """``fields.py:146-148`` is_fixed as a bool array."""
    m = np.zeros(n_dofs, dtype=bool)
    m[np.asarray(fixed_dofs, dtype=np.int64)] = True
    return m


def scale(rho: np.ndarray, E_min: float, penal: float) -> np.ndarray:
    """SIMP interpolation ``E_min + rho^p (1 - E_min)`` (``fem/kernels.py:92,160,268``)."""
    return E_min + np.power(rho, penal) * (1.0 - E_min)


# --------------------------------------------------------------------------
# FEM operators
# --------------------------------------------------------------------------


def matvec(p, s, Ke, edof, fixed) -> np.ndarray:
    """K(rho) p by element scatter; fixed rows become identity, fixed columns
    are kept (``fem/kernels.py:143-176``)."""
    contrib = s[:, None] * (p[edof] @ Ke.T)
    Ap = np.bincount(edof.ravel(), weights=contrib.ravel(), minlength=p.size)
    Ap[fixed] = p[fixed]
    return Ap


def diagonal(s, Ke, edof, fixed) -> np.ndarray:
    """diag K(rho); fixed -> 1; clamp 1e-20 (``fem/kernels.py:76-101``)."""
    contrib = s[:, None] * np.diag(Ke)[None, :]
    d = np.bincount(edof.ravel(), weights=contrib.ravel(), minlength=fixed.size)
    d[fixed] = 1.0
    d[(~fixed) & (d < 1e-20)] = 1e-20
    return d


def ke_inverse_regularised(Ke):
    """``fields.py:139-143``: inv(Ke + alpha I), alpha = 0.1 mean|diag Ke|."""
    alpha = 0.1 * float(np.mean(np.abs(np.diag(Ke))))
    return np.linalg.inv(Ke + alpha * np.eye(Ke.shape[0])), alpha


def dof_count(edof, n_dofs) -> np.ndarray:
    """Elements sharing each DOF (``fem/preconditioner.py:19-30``)."""
    return np.bincount(edof.ravel(), minlength=n_dofs)


def elemblock_apply(r, s, Ke_inv, alpha_reg, edof, count, fixed) -> np.ndarray:
    """Symmetric element-block Jacobi (``fem/preconditioner.py:47-108``):
    ``z = C^-1/2 sum_e w_e P_e^T Ke_inv P_e C^-1/2 r``, ``w_e = s/max(s,a)^2``,
    fixed z = 0."""
    sc = np.maximum(s, alpha_reg)
    w = s / (sc * sc)
    root = np.sqrt(count.astype(np.float64))
    zl = (r[edof] / root[edof]) @ Ke_inv.T
    z = np.bincount(edof.ravel(), weights=(w[:, None] * zl).ravel(), minlength=r.size)
    z = z / root
    z[fixed] = 0.0
    return z


def block2x2_inverse(s, Ke, edof, fixed):
    """Nodal 2x2 block-Jacobi inverse (no reference counterpart).

    ``D_n = sum_{e ni n} s_e Ke[2a:2a+2, 2a:2a+2]`` (a = local node of n in e);
    a fixed DOF gets an identity row and column. Returns (n_nodes, 2, 2).
    """
    n_nodes = fixed.size // 2
    D = np.zeros((n_nodes, 2, 2))
    for a in range(4):
        nodes = edof[:, 2 * a] // 2
        blk = s[:, None, None] * Ke[2 * a:2 * a + 2, 2 * a:2 * a + 2][None]
        for i in range(2):
            for j in range(2):
                D[:, i, j] += np.bincount(nodes, weights=blk[:, i, j], minlength=n_nodes)
    fx = fixed.reshape(-1, 2)
    for c in range(2):
        D[fx[:, c], c, :] = 0.0
        D[fx[:, c], :, c] = 0.0
        D[fx[:, c], c, c] = 1.0
    return np.linalg.inv(D)


def block2x2_apply(r, Dinv) -> np.ndarray:
    return np.einsum("nij,nj->ni", Dinv, r.reshape(-1, 2)).ravel()


def pcg(f, s, Ke, edof, fixed, u_prev=None, warm_start=False, max_iter=2000, tol=1e-8,
        preconditioner="jacobi", return_history=False):
    """Matrix-free PCG restating ``fem/solver.py:27-159`` in FP64.

    Quirks kept: convergence on the recursively updated ``|r|^2 < (tol |f|)^2``;
    at least one iteration; ``|pAp| < 1e-30`` breaks and reports ``max_iter``;
    ``|f| < 1e-30`` gives u = 0 and 0 iterations; ``beta = rz/(rz_old + 1e-60)``.
    Returns (u, n_iter[, history of |r|^2]).
    """
    n = f.size
    f_norm = math.sqrt(float(np.dot(f, f)))
    if f_norm < 1e-30:
        out = (np.zeros(n), 0)
        return out + ([],) if return_history else out
    tol_sq = (tol * f_norm) ** 2

    if preconditioner == "jacobi":
        dinv = diagonal(s, Ke, edof, fixed)
        apply = lambda r: r / dinv  # noqa: E731  (solver.py:91, fused_kernels.py:125)
    elif preconditioner == "block_jacobi":
        Ke_inv, a_reg = ke_inverse_regularised(Ke)
        cnt = dof_count(edof, n)
        apply = lambda r: elemblock_apply(r, s, Ke_inv, a_reg, edof, cnt, fixed)  # noqa: E731
    elif preconditioner == "block_jacobi_2x2":
        Dinv = block2x2_inverse(s, Ke, edof, fixed)
        apply = lambda r: block2x2_apply(r, Dinv)  # noqa: E731
    elif preconditioner == "none":
        apply = lambda r: r.copy()  # noqa: E731
    else:
        raise ValueError(preconditioner)

    u = u_prev.astype(np.float64).copy() if (warm_start and u_prev is not None) else np.zeros(n)
    u[fixed] = 0.0
    r = f - matvec(u, s, Ke, edof, fixed)
    z = apply(r)
    p = z.copy()
    rz_old = float(np.dot(r, z))
    hist = []
    for it in range(max_iter):
        Ap = matvec(p, s, Ke, edof, fixed)
        pAp = float(np.dot(p, Ap))
        if abs(pAp) < 1e-30:
            break
        alpha = rz_old / pAp
        u += alpha * p
        r -= alpha * Ap
        z = apply(r)
        rz_new = float(np.dot(r, z))
        rr = float(np.dot(r, r))
        hist.append(rr)
        if rr < tol_sq:
            return (u, it + 1, hist) if return_history else (u, it + 1)
        beta = rz_new / (rz_old + 1e-60)
        p = z + beta * p
        rz_old = rz_new
    return (u, max_iter, hist) if return_history else (u, max_iter)


def sensitivity(u, rho, Ke, edof, E_min, penal):
    """Compliance and sensitivities (``fem/kernels.py:234-272``):
    ``dc = -p rho^(p-1) (1-E_min) u_e^T Ke u_e``, ``dv = 1``,
    ``c = sum s_e u_e^T Ke u_e``."""
    ue = u[edof]
    uKu = np.einsum("ei,ij,ej->e", ue, Ke, ue)
    dc = -penal * np.power(rho, penal - 1.0) * (1.0 - E_min) * uKu
    c = float(np.sum(scale(rho, E_min, penal) * uKu))
    return c, dc, np.ones_like(rho)


# --------------------------------------------------------------------------
# filter
# --------------------------------------------------------------------------


def filter_tables_bruteforce(nelx: int, nely: int, rmin: float):
    """Padded neighbour tables exactly as ``optimizer/filter.py:22-74,99-103``.

    Pure Python loops with the reference's arithmetic: float32 centres
    (``mesh.py:123-136``), distances in float32 (NumPy 2 scalar rules),
    ``math.sqrt`` in FP64, weights stored as float32. Small grids only.
    """
    r_ceil = math.ceil(rmin)
    width = (2 * r_ceil + 1) ** 2
    n_elem = nelx * nely
    nb = np.full((n_elem, width), -1, dtype=np.int32)
    wt = np.zeros((n_elem, width), dtype=np.float32)
    cnt = np.zeros(n_elem, dtype=np.int32)
    for ex in range(nelx):
        for ey in range(nely):
            e = ex * nely + ey
            cx, cy = np.float32(ex + 0.5), np.float32(ey + 0.5)
            k = 0
            for jx in range(max(0, ex - r_ceil), min(nelx, ex + r_ceil + 1)):
                for jy in range(max(0, ey - r_ceil), min(nely, ey + r_ceil + 1)):
                    dist = math.sqrt((cx - (jx + 0.5)) ** 2 + (cy - (jy + 0.5)) ** 2)
                    w = max(0.0, rmin - dist)
                    if w > 0.0:
                        nb[e, k] = jx * nely + jy
                        wt[e, k] = w
                        k += 1
            cnt[e] = k
    return nb, wt, cnt


def filter_stencil(rmin: float):
    """Compact stencil (dx, dy, w_f32) in the reference's table order
    (jx outer, jy inner; ``filter.py:64-72``). Weights rounded to float32."""
    r_ceil = math.ceil(rmin)
    out = []
    for dx in range(-r_ceil, r_ceil + 1):
        for dy in range(-r_ceil, r_ceil + 1):
            w = max(0.0, rmin - math.sqrt(float(dx * dx + dy * dy)))
            if w > 0.0:
                out.append((dx, dy, np.float32(w)))
    return out


def filter_tables(nelx: int, nely: int, rmin: float):
    """Vectorised expansion of the stencil into the reference's padded tables.
    Equality with ``filter_tables_bruteforce`` is a test."""
    st = filter_stencil(rmin)
    width = (2 * math.ceil(rmin) + 1) ** 2
    n_elem = nelx * nely
    ex, ey = np.divmod(np.arange(n_elem), nely)
    nb = np.full((n_elem, width), -1, dtype=np.int32)
    wt = np.zeros((n_elem, width), dtype=np.float32)
    cnt = np.zeros(n_elem, dtype=np.int32)
    rows = np.arange(n_elem)
    for dx, dy, w in st:
        jx, jy = ex + dx, ey + dy
        ok = (jx >= 0) & (jx < nelx) & (jy >= 0) & (jy < nely)
        nb[rows[ok], cnt[ok]] = (jx * nely + jy)[ok]
        wt[rows[ok], cnt[ok]] = w
        cnt[ok] += 1
    return nb, wt, cnt


def filter_tables_bruteforce_3d(nelx: int, nely: int, nelz: int, rmin: float):
    """3-D padded tables exactly as ``optimizer/filter.py:75-97,99-103``
    (float32 centres, FP64 sqrt, float32 weights; jx outer, jy, jz inner)."""
    r_ceil = math.ceil(rmin)
    width = (2 * r_ceil + 1) ** 3
    n_elem = nelx * nely * nelz
    nb = np.full((n_elem, width), -1, dtype=np.int32)
    wt = np.zeros((n_elem, width), dtype=np.float32)
    cnt = np.zeros(n_elem, dtype=np.int32)
    for ex in range(nelx):
        for ey in range(nely):
            for ez in range(nelz):
                e = ex * nely * nelz + ey * nelz + ez
                cx, cy, cz = np.float32(ex + 0.5), np.float32(ey + 0.5), np.float32(ez + 0.5)
                k = 0
                for jx in range(max(0, ex - r_ceil), min(nelx, ex + r_ceil + 1)):
                    for jy in range(max(0, ey - r_ceil), min(nely, ey + r_ceil + 1)):
                        for jz in range(max(0, ez - r_ceil), min(nelz, ez + r_ceil + 1)):
                            dist = math.sqrt((cx - (jx + 0.5)) ** 2 + (cy - (jy + 0.5)) ** 2
                                             + (cz - (jz + 0.5)) ** 2)
                            w = max(0.0, rmin - dist)
                            if w > 0.0:
                                nb[e, k] = jx * nely * nelz + jy * nelz + jz
                                wt[e, k] = w
                                k += 1
                cnt[e] = k
    return nb, wt, cnt


def filter_tables_3d(nelx: int, nely: int, nelz: int, rmin: float):
    """Vectorised 3-D tables from the stencil in table order; equality with
    ``filter_tables_bruteforce_3d`` is a test."""
    r_ceil = math.ceil(rmin)
    width = (2 * r_ceil + 1) ** 3
    n_elem = nelx * nely * nelz
    e = np.arange(n_elem)
    ex, rem = np.divmod(e, nely * nelz)
    ey, ez = np.divmod(rem, nelz)
    nb = np.full((n_elem, width), -1, dtype=np.int32)
    wt = np.zeros((n_elem, width), dtype=np.float32)
    cnt = np.zeros(n_elem, dtype=np.int32)
    for dx in range(-r_ceil, r_ceil + 1):
        for dy in range(-r_ceil, r_ceil + 1):
            for dz in range(-r_ceil, r_ceil + 1):
                w = max(0.0, rmin - math.sqrt(float(dx * dx + dy * dy + dz * dz)))
                if w <= 0.0:
                    continue
                jx, jy, jz = ex + dx, ey + dy, ez + dz
                ok = (jx >= 0) & (jx < nelx) & (jy >= 0) & (jy < nely) & (jz >= 0) & (jz < nelz)
                nb[e[ok], cnt[ok]] = ((jx * nely + jy) * nelz + jz)[ok]
                wt[e[ok], cnt[ok]] = np.float32(w)
                cnt[ok] += 1
    return nb, wt, cnt


def _gather_sum(vals, nb, wt, cnt):
    """Sequential table-order sums sum_k w_k vals[nb_k] and sum_k w_k (FP64)."""
    num = np.zeros(nb.shape[0])
    den = np.zeros(nb.shape[0])
    for k in range(nb.shape[1]):
        act = k < cnt
        if not act.any():
            break
        w = wt[:, k].astype(np.float64) * act
        num += w * vals[np.where(act, nb[:, k], 0)]
        den += w
    return num, den


def density_filter(rho, nb, wt, cnt):
    """``filter.py:106-124``: ``H rho / (sum H + 1e-20)``."""
    num, den = _gather_sum(rho, nb, wt, cnt)
    return num / (den + 1e-20)


def dc_average_filter(dc, nb, wt, cnt):
    """``filter.py:127-149``: ``H dc / (sum H + 1e-20)`` (the reference's
    density-mode sensitivity projection; not the exact chain rule)."""
    num, den = _gather_sum(dc, nb, wt, cnt)
    return num / (den + 1e-20)


def sensitivity_filter(rho, dc, nb, wt, cnt):
    """``filter.py:152-171``: ``H (rho dc) / (rho_e sum H + 1e-20)``."""
    num, den = _gather_sum(rho * dc, nb, wt, cnt)
    return num / (rho * den + 1e-20)


def filter_transpose(g, nb, wt, cnt):
    """Exact chain rule of the density filter (no reference counterpart):
    ``H^T (g / (Hs + 1e-20))``, with H as the reference stores it."""
    _, den = _gather_sum(np.zeros_like(g), nb, wt, cnt)
    y = g / (den + 1e-20)
    out = np.zeros_like(g)
    for k in range(nb.shape[1]):
        act = k < cnt
        if not act.any():
            break
        np.add.at(out, nb[act, k], wt[act, k].astype(np.float64) * y[act])
    return out


# --------------------------------------------------------------------------
# optimality criteria
# --------------------------------------------------------------------------


def oc_candidate(rho, dc, dv, lam, move):
    """Per-element OC update for one lambda (``optimizer/oc.py:40-47``)."""
    be = np.maximum(-dc / (lam * dv + 1e-20), 0.0)
    cand = rho * np.sqrt(be)
    lo = np.maximum(0.001, rho - move)
    hi = np.minimum(1.0, rho + move)
    return np.maximum(lo, np.minimum(hi, cand))


def oc_update(rho, dc, dv, volfrac, move, return_trace=False):
    """Bisection on lambda (``optimizer/oc.py:52-93``): bracket [0, 1e9], at
    most 50 steps, stop when (hi-lo)/(hi+lo+1e-30) < 1e-4, commit the rho of
    the LAST lambda evaluated. Returns (rho_new, max_change[, trace])."""
    lo, hi = 0.0, 1e9
    change = 0.0
    rho_new = rho
    trace = []
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        rho_new = oc_candidate(rho, dc, dv, mid, move)
        vol = float(np.sum(rho_new)) / rho.size
        change = float(np.max(np.abs(rho_new - rho)))
        trace.append((mid, vol))
        if vol > volfrac:
            lo = mid
        else:
            hi = mid
        if (hi - lo) / (hi + lo + 1e-30) < 1e-4:
            break
    return (rho_new, change, trace) if return_trace else (rho_new, change)


# --------------------------------------------------------------------------
# SIMP loop
# --------------------------------------------------------------------------


def simp_run(nelx, nely, fixed_dofs, force_dofs, force_values, *, E=1.0, nu=0.3, E_min=1e-9,
             volfrac=0.5, penal=3.0, rmin=1.5, filter_type="density", max_iter=50, tol=0.01,
             move=0.2, max_cg=2000, cg_tol=1e-8, preconditioner="jacobi"):
    """The optimization loop of ``optimizer/simp.py:99-219`` (density mode:
    filter -> solve on filtered rho -> sensitivity on filtered rho -> H dc/Hs
    -> OC; sensitivity mode: solve and sensitivity on rho, Sigmund filter).

    Returns a dict with per-iteration lists ``c, vol, change, n_cg`` and the
    final ``rho`` (flat, e = ex*nely+ey) and ``u``.
    """
    n_dofs = 2 * (nelx + 1) * (nely + 1)
    return _simp_loop(ke_q4(E, nu), edof_q4(nelx, nely), n_dofs, filter_tables(nelx, nely, rmin),
                      fixed_dofs, force_dofs, force_values, E_min=E_min, volfrac=volfrac, penal=penal,
                      filter_type=filter_type, max_iter=max_iter, tol=tol, move=move, max_cg=max_cg,
                      cg_tol=cg_tol, preconditioner=preconditioner)


def simp_run_3d(nelx, nely, nelz, fixed_dofs, force_dofs, force_values, *, E=1.0, nu=0.3, E_min=1e-9,
                volfrac=0.5, penal=3.0, rmin=1.5, filter_type="density", max_iter=50, tol=0.01,
                move=0.2, max_cg=2000, cg_tol=1e-8, preconditioner="jacobi"):
    """The same loop on the 3-D H8 grid (``optimizer/simp.py:99-219`` with
    dim = 3): H8 Ke, 3-D connectivity and 3-D filter tables; rho flat in
    e = (ex*nely + ey)*nelz + ez order."""
    n_dofs = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    return _simp_loop(ke_h8(E, nu), edof_h8(nelx, nely, nelz), n_dofs, filter_tables_3d(nelx, nely, nelz, rmin),
                      fixed_dofs, force_dofs, force_values, E_min=E_min, volfrac=volfrac, penal=penal,
                      filter_type=filter_type, max_iter=max_iter, tol=tol, move=move, max_cg=max_cg,
                      cg_tol=cg_tol, preconditioner=preconditioner)


def _simp_loop(Ke, edof, n_dofs, tables, fixed_dofs, force_dofs, force_values, *, E_min, volfrac, penal,
               filter_type, max_iter, tol, move, max_cg, cg_tol, preconditioner):
    n_elem = edof.shape[0]
    fixed = fixed_mask(n_dofs, fixed_dofs)
    f = force_vector(n_dofs, force_dofs, force_values, fixed_dofs)
    nb, wt, cnt = tables
    rho = np.full(n_elem, float(volfrac))
    u_prev = np.zeros(n_dofs)
    hist = {"c": [], "vol": [], "change": [], "n_cg": []}
    u = u_prev
    for it in range(1, max_iter + 1):
        if filter_type == "density":
            rho_phys = density_filter(rho, nb, wt, cnt)
        else:
            rho_phys = rho
        s = scale(rho_phys, E_min, penal)
        u, n_cg = pcg(f, s, Ke, edof, fixed, u_prev, warm_start=(it > 1), max_iter=max_cg,
                      tol=cg_tol, preconditioner=preconditioner)
        if n_cg > 0 or np.dot(f, f) >= 1e-60:
            u_prev = u
        c, dc, dv = sensitivity(u, rho_phys, Ke, edof, E_min, penal)
        if filter_type == "density":
            dc_f = dc_average_filter(dc, nb, wt, cnt)
        else:
            dc_f = sensitivity_filter(rho, dc, nb, wt, cnt)
        rho, change = oc_update(rho, dc_f, dv, volfrac, move)
        vol = float(np.sum(rho)) / rho.size
        for k, v in (("c", c), ("vol", vol), ("change", change), ("n_cg", n_cg)):
            hist[k].append(v)
        if it > 5 and change < tol:
            break
    hist["rho"] = rho
    hist["u"] = u
    return hist
