/* Body of the C restatement of the reference SIMP hot path (TEST
 * INFRASTRUCTURE), instantiated twice by oracle.c: REAL = double (suffix
 * _f64, the oracle for large grids) and REAL = float (suffix _f32, the
 * reference's own precision, used as the timed CPU baseline).
 *
 * Follows (paths relative to the reference src/toptimization/):
 *   material.py:20-86 (Ke, closed form), mesh.py:33-65 (edof numbering),
 *   fem/kernels.py:76-101,143-176,234-272,330-336, fem/solver.py:27-159,
 *   optimizer/filter.py:22-74,106-171, optimizer/oc.py:21-93,
 *   optimizer/simp.py:99-219.
 * The element scatter of fem/kernels.py:159-172 uses atomics; here elements
 * are processed in four colours ((ex%2, ey%2): same-colour elements share no
 * node), which is the same sum without atomics and is deterministic.
 */

static void SUF(ke)(double E, double nu, REAL* Ke) {
  const double k[8] = {0.5 - nu / 6, 0.125 + nu / 8, -0.25 - nu / 12, -0.125 + 3 * nu / 8,
                       -0.25 + nu / 12, -0.125 - nu / 8, nu / 6, 0.125 - 3 * nu / 8};
  static const int idx[8][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {1, 0, 7, 6, 5, 4, 3, 2}, {2, 7, 0, 5, 6, 3, 4, 1},
                                {3, 6, 5, 0, 7, 2, 1, 4}, {4, 5, 6, 7, 0, 1, 2, 3}, {5, 4, 3, 2, 1, 0, 7, 6},
                                {6, 3, 4, 1, 2, 7, 0, 5}, {7, 2, 1, 4, 3, 6, 5, 0}};
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) Ke[i * 8 + j] = (REAL)(E / (1.0 - nu * nu) * k[idx[i][j]]);
}

static inline void SUF(edof)(int nely, int ex, int ey, long* d) {
  const long n1 = (long)ex * (nely + 1) + ey, n2 = (long)(ex + 1) * (nely + 1) + ey;
  d[0] = 2 * n1; d[1] = 2 * n1 + 1; d[2] = 2 * n2; d[3] = 2 * n2 + 1;
  d[4] = 2 * n2 + 2; d[5] = 2 * n2 + 3; d[6] = 2 * n1 + 2; d[7] = 2 * n1 + 3;
}

/* Ap = K(s) p; fixed rows -> Ap = p (fem/kernels.py:143-176). */
void SUF(orc_matvec)(int nelx, int nely, const REAL* Ke, const REAL* s, const unsigned char* fixed,
                     const REAL* p, REAL* Ap) {
  const long nd = 2L * (nelx + 1) * (nely + 1);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nd; ++i) Ap[i] = 0;
  for (int color = 0; color < 4; ++color) {
#pragma omp parallel for schedule(static)
    for (int ex = color >> 1; ex < nelx; ex += 2) {
      for (int ey = color & 1; ey < nely; ey += 2) {
        long d[8];
        SUF(edof)(nely, ex, ey, d);
        const REAL se = s[(long)ex * nely + ey];
        REAL pl[8];
        for (int i = 0; i < 8; ++i) pl[i] = p[d[i]];
        for (int i = 0; i < 8; ++i) {
          REAL acc = 0;
          for (int j = 0; j < 8; ++j) acc += Ke[i * 8 + j] * pl[j];
          Ap[d[i]] += se * acc;
        }
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nd; ++i)
    if (fixed[i]) Ap[i] = p[i];
}

static void SUF(diagonal)(int nelx, int nely, const REAL* Ke, const REAL* s, const unsigned char* fixed,
                          REAL* diag) {
  const long nd = 2L * (nelx + 1) * (nely + 1);
  for (long i = 0; i < nd; ++i) diag[i] = 0;
  for (int color = 0; color < 4; ++color) {
#pragma omp parallel for schedule(static)
    for (int ex = color >> 1; ex < nelx; ex += 2)
      for (int ey = color & 1; ey < nely; ey += 2) {
        long d[8];
        SUF(edof)(nely, ex, ey, d);
        const REAL se = s[(long)ex * nely + ey];
        for (int i = 0; i < 8; ++i) diag[d[i]] += se * Ke[i * 9];
      }
  }
  for (long i = 0; i < nd; ++i) {
    if (fixed[i]) diag[i] = 1;
    else if (diag[i] < (REAL)1e-20) diag[i] = (REAL)1e-20;
  }
}

static double SUF(dot)(long n, const REAL* a, const REAL* b) {
  REAL acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
  for (long i = 0; i < n; ++i) acc += a[i] * b[i];
  return (double)acc;
}

/* PCG of fem/solver.py:27-159 (Jacobi when jacobi != 0, else none).
 * work: 5 * nd REALs. Returns the reference's iteration count. */
int SUF(orc_pcg)(int nelx, int nely, const REAL* Ke, const REAL* s, const unsigned char* fixed,
                 const REAL* f, REAL* u, int warm, int max_iter, double tol, int jacobi, REAL* work) {
  const long nd = 2L * (nelx + 1) * (nely + 1);
  REAL *r = work, *p = work + nd, *z = work + 2 * nd, *Ap = work + 3 * nd, *diag = work + 4 * nd;
  const double f_norm = sqrt(SUF(dot)(nd, f, f));
  if (f_norm < 1e-30) {
    for (long i = 0; i < nd; ++i) u[i] = 0;
    return 0;
  }
  const double tol_sq = (tol * f_norm) * (tol * f_norm);
  if (jacobi) SUF(diagonal)(nelx, nely, Ke, s, fixed, diag);
  if (!warm)
    for (long i = 0; i < nd; ++i) u[i] = 0;
  for (long i = 0; i < nd; ++i)
    if (fixed[i]) u[i] = 0;
  SUF(orc_matvec)(nelx, nely, Ke, s, fixed, u, Ap);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nd; ++i) {
    r[i] = f[i] - Ap[i];
    z[i] = jacobi ? r[i] / diag[i] : r[i];
    p[i] = z[i];
  }
  double rz_old = SUF(dot)(nd, r, z);
  for (int it = 0; it < max_iter; ++it) {
    SUF(orc_matvec)(nelx, nely, Ke, s, fixed, p, Ap);
    const double pAp = SUF(dot)(nd, p, Ap);
    if (fabs(pAp) < 1e-30) break;
    const REAL alpha = (REAL)(rz_old / pAp);
    REAL rz = 0, rr = 0;
#pragma omp parallel for reduction(+ : rz, rr) schedule(static)
    for (long i = 0; i < nd; ++i) {
      u[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      z[i] = jacobi ? r[i] / diag[i] : r[i];
      rz += r[i] * z[i];
      rr += r[i] * r[i];
    }
    if ((double)rr < tol_sq) return it + 1;
    const REAL beta = (REAL)((double)rz / (rz_old + 1e-60));
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nd; ++i) p[i] = z[i] + beta * p[i];
    rz_old = (double)rz;
  }
  return max_iter;
}

/* Compliance + dc (fem/kernels.py:234-272). */
double SUF(orc_sensitivity)(int nelx, int nely, const REAL* Ke, const REAL* u, const REAL* rho, double E_min,
                            double penal, REAL* dc) {
  REAL c = 0;
#pragma omp parallel for reduction(+ : c) schedule(static)
  for (long e = 0; e < (long)nelx * nely; ++e) {
    const int ex = (int)(e / nely), ey = (int)(e % nely);
    long d[8];
    SUF(edof)(nely, ex, ey, d);
    REAL ul[8], uKu = 0;
    for (int i = 0; i < 8; ++i) ul[i] = u[d[i]];
    for (int i = 0; i < 8; ++i) {
      REAL t = 0;
      for (int j = 0; j < 8; ++j) t += Ke[i * 8 + j] * ul[j];
      uKu += ul[i] * t;
    }
    const REAL re = rho[e];
    const REAL se = (REAL)E_min + (REAL)pow(re, penal) * (REAL)(1.0 - E_min);
    dc[e] = (REAL)(-penal) * (REAL)pow(re, penal - 1.0) * (REAL)(1.0 - E_min) * uKu;
    c += se * uKu;
  }
  return (double)c;
}

/* Cone filter in table order (filter.py:58-74, 106-171). mode 0 density,
 * 1 dc average, 2 sensitivity (Sigmund). Weights rounded to float as the
 * reference stores them. */
void SUF(orc_filter)(int nelx, int nely, double rmin, int mode, const REAL* rho, const REAL* dc, REAL* out) {
  const int rc = (int)ceil(rmin);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < (long)nelx * nely; ++e) {
    const int ex = (int)(e / nely), ey = (int)(e % nely);
    REAL num = 0, den = 0;
    const int x0 = ex - rc > 0 ? ex - rc : 0, x1 = ex + rc + 1 < nelx ? ex + rc + 1 : nelx;
    const int y0 = ey - rc > 0 ? ey - rc : 0, y1 = ey + rc + 1 < nely ? ey + rc + 1 : nely;
    for (int jx = x0; jx < x1; ++jx)
      for (int jy = y0; jy < y1; ++jy) {
        const double w64 = rmin - sqrt((double)((ex - jx) * (ex - jx) + (ey - jy) * (ey - jy)));
        if (!(w64 > 0.0)) continue;
        const REAL w = (REAL)(float)w64;
        const long j = (long)jx * nely + jy;
        const REAL v = mode == 0 ? rho[j] : (mode == 1 ? dc[j] : rho[j] * dc[j]);
        num += w * v;
        den += w;
      }
    out[e] = mode == 2 ? num / (rho[e] * den + (REAL)1e-20) : num / (den + (REAL)1e-20);
  }
}

/* OC bisection (oc.py:52-93); rho updated in place; returns max change. */
double SUF(orc_oc)(long ne, REAL* rho, const REAL* dcf, double volfrac, double move, REAL* tmp, int* steps_out,
                   double* lam_out) {
  double lo = 0.0, hi = 1e9, change = 0.0, mid = 0.0;
  int steps = 0;
  for (int k = 0; k < 50; ++k) {
    mid = 0.5 * (lo + hi);
    const REAL lam = (REAL)mid;
    REAL vol = 0, mx = 0;
#pragma omp parallel for reduction(+ : vol) reduction(max : mx) schedule(static)
    for (long e = 0; e < ne; ++e) {
      REAL be = -dcf[e] / (lam * (REAL)1 + (REAL)1e-20);
      be = be > 0 ? be : 0;
      const REAL cand = rho[e] * (REAL)sqrt(be);
      REAL lo_b = rho[e] - (REAL)move, hi_b = rho[e] + (REAL)move;
      lo_b = lo_b > (REAL)0.001 ? lo_b : (REAL)0.001;
      hi_b = hi_b < (REAL)1 ? hi_b : (REAL)1;
      const REAL c = cand < hi_b ? cand : hi_b;
      tmp[e] = c > lo_b ? c : lo_b;
      vol += tmp[e];
      const REAL d = (REAL)fabs((double)(tmp[e] - rho[e]));
      mx = d > mx ? d : mx;
    }
    change = (double)mx;
    ++steps;
    if ((double)vol / (double)ne > volfrac) lo = mid;
    else hi = mid;
    if ((hi - lo) / (hi + lo + 1e-30) < 1e-4) break;
  }
  for (long e = 0; e < ne; ++e) rho[e] = tmp[e];
  if (steps_out) *steps_out = steps;
  if (lam_out) *lam_out = mid;
  return change;
}

/* SIMP loop, density (filter_type 0) or sensitivity (1) mode
 * (simp.py:99-219). Per-iteration outputs c, vol, change, ncg (length
 * max_iter); returns the number of iterations run. rho (in: ignored, out:
 * final), u (out: final). Times the loop into *seconds. */
int SUF(orc_simp)(int nelx, int nely, double E, double nu, double E_min, double penal, double volfrac,
                  double rmin, int filter_type, int max_iter, double tol, double move, int max_cg, double cg_tol,
                  int jacobi, const unsigned char* fixed, const double* f_in, REAL* rho, REAL* u, double* c_out,
                  double* vol_out, double* change_out, int* ncg_out, double* seconds) {
  const long nd = 2L * (nelx + 1) * (nely + 1), ne = (long)nelx * nely;
  REAL Ke[64];
  SUF(ke)(E, nu, Ke);
  REAL* f = (REAL*)malloc(sizeof(REAL) * nd);
  REAL* work = (REAL*)malloc(sizeof(REAL) * 5 * nd);
  REAL* rf = (REAL*)malloc(sizeof(REAL) * ne);
  REAL* s = (REAL*)malloc(sizeof(REAL) * ne);
  REAL* dc = (REAL*)malloc(sizeof(REAL) * ne);
  REAL* dcf = (REAL*)malloc(sizeof(REAL) * ne);
  REAL* tmp = (REAL*)malloc(sizeof(REAL) * ne);
  for (long i = 0; i < nd; ++i) f[i] = fixed[i] ? 0 : (REAL)f_in[i];
  for (long e = 0; e < ne; ++e) rho[e] = (REAL)volfrac;
  for (long i = 0; i < nd; ++i) u[i] = 0;
  const double t0 = omp_get_wtime();
  int it = 1;
  for (; it <= max_iter; ++it) {
    const REAL* phys = rho;
    if (filter_type == 0) {
      SUF(orc_filter)(nelx, nely, rmin, 0, rho, NULL, rf);
      phys = rf;
    }
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) s[e] = (REAL)E_min + (REAL)pow(phys[e], penal) * (REAL)(1.0 - E_min);
    const int ncg = SUF(orc_pcg)(nelx, nely, Ke, s, fixed, f, u, it > 1, max_cg, cg_tol, jacobi, work);
    const double c = SUF(orc_sensitivity)(nelx, nely, Ke, u, phys, E_min, penal, dc);
    if (filter_type == 0) SUF(orc_filter)(nelx, nely, rmin, 1, NULL, dc, dcf);
    else SUF(orc_filter)(nelx, nely, rmin, 2, rho, dc, dcf);
    const double change = SUF(orc_oc)(ne, rho, dcf, volfrac, move, tmp, NULL, NULL);
    REAL vs = 0;
    for (long e = 0; e < ne; ++e) vs += rho[e];
    c_out[it - 1] = c;
    vol_out[it - 1] = (double)vs / (double)ne;
    change_out[it - 1] = change;
    ncg_out[it - 1] = ncg;
    if (it > 5 && change < tol) break;
  }
  if (seconds) *seconds = omp_get_wtime() - t0;
  free(f); free(work); free(rf); free(s); free(dc); free(dcf); free(tmp);
  return it > max_iter ? max_iter : it;
}
