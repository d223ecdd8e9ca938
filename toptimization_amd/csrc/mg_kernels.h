// Geometric multigrid preconditioner for the PCG (PC_MG). No reference
// counterpart: the reference offers Jacobi and element-block Jacobi only
// (fem/preconditioner.py); this is an additional preconditioner for the
// same solve (fem/solver.py:27-167), z = M r with M one V(1,1) cycle.
//
// Hierarchy. Level l+1 agglomerates 2x2 patches of level-l elements; the
// prolongation P is bilinear interpolation of the coarse nodes (coarse node
// (I, J) sits on fine node (2I, 2J)); at level 0 it is masked by the fixed
// DOFs (fixed DOFs receive no correction). Coarse operators are Galerkin
// products K_{l+1} = P^T K_l P, formed element by element: a coarse element
// matrix is sum over its four children c of P_c^T K_c P_c, with P_c the 8x8
// interpolation from the coarse element's corners to child c's corners and
// K_c = s_c Ke (fixed rows and columns zeroed) at level 0. The coarsest level
// (<= TOPT_MG_DENSE DOFs) is solved with a dense inverse formed once per
// solve. Smoother: damped Jacobi (omega 0.6), one sweep before and one after
// the coarse correction, so M is symmetric positive definite and plain PCG
// applies. The iteration count depends little on the grid size (a 300x100
// bridge after 20 OC iterations: 69 iterations against 1,738 with Jacobi).
#pragma once
#include "fem_kernels.h"

namespace topt {

#define TOPT_MG_DENSE 128      // max DOFs of the coarsest level (inverted in LDS, FP64: 128 KB)
// Jacobi damping. Every element matrix (s_e Ke, with or without fixed-DOF
// masking, and every Galerkin agglomerate of them) has lambda_max(D_e^-1
// K_e) = 2.889 for the unit-square Q4 with nu = 0.3, so lambda_max(D^-1 K)
// <= 2.889 on every level whatever the densities, and a damped Jacobi step
// converges in the energy norm for omega < 2 / 2.889 = 0.69. 0.6 keeps the
// V-cycle SPD with margin (0.7 and 0.8 save 4-8 % of the iterations on
// sample designs but leave the guaranteed range; 0.9 diverges).
#define TOPT_MG_OMEGA 0.6

// Packed index of entry (i, j) of a symmetric 8x8 matrix (upper triangle,
// row major: row i holds columns i..7).
__host__ __device__ constexpr int sym8(int i, int j) {
  return i <= j ? i * 8 - i * (i - 1) / 2 + (j - i) : j * 8 - j * (j - 1) / 2 + (i - j);
}
// Local node k of an element: BL, BR, TR, TL (mesh.py:17-24) at offsets (LX, LY).
__host__ __device__ constexpr int mg_lx(int k) { return (k == 1 || k == 2) ? 1 : 0; }
__host__ __device__ constexpr int mg_ly(int k) { return (k >= 2) ? 1 : 0; }
// Weight of coarse corner k at child c = (cx, cy)'s corner a (positions in
// fine element units inside the coarse element [0, 2]^2).
__host__ __device__ constexpr double mg_w(int cx, int cy, int a, int k) {
  return (mg_lx(k) ? 0.5 * (cx + mg_lx(a)) : 1.0 - 0.5 * (cx + mg_lx(a))) *
         (mg_ly(k) ? 0.5 * (cy + mg_ly(a)) : 1.0 - 0.5 * (cy + mg_ly(a)));
}
// Local index of node (ix, iy) in the element at (ix - 1 + a, iy - 1 + b).
__host__ __device__ constexpr int mg_me(int a, int b) { return (a == 1) ? ((b == 1) ? 0 : 3) : ((b == 1) ? 1 : 2); }

// ---------------------------------------------------------------- setup
// Coarse element matrices of level l+1 from level l (FINE: l = 0, children
// s_e Ke with fixed DOFs masked). One thread per coarse element, FP64 sums.
template <typename T, bool FINE, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_galerkin(const ProbDev* tab, const int2* map, ProbDev single,
                                                           int l) {
  int lb;
  PD& P = grid_prob<B>(tab, map, single, blockIdx.x, lb);
  const int fnx = P.mg[l].nx, fny = P.mg[l].ny;
  const int cnx = P.mg[l + 1].nx, cny = P.mg[l + 1].ny;
  const long E = (long)lb * TOPT_BLOCK + threadIdx.x;
  if (E >= (long)cnx * cny) return;
  const int I = (int)(E / cny), J = (int)(E - (long)I * cny);
  double kc[36];
#pragma unroll
  for (int t = 0; t < 36; ++t) kc[t] = 0.0;
#pragma unroll
  for (int cx = 0; cx < 2; ++cx)
#pragma unroll
    for (int cy = 0; cy < 2; ++cy) {
      const int ex = 2 * I + cx, ey = 2 * J + cy;
      if (ex >= fnx || ey >= fny) continue;   // padding: void
      double kf[36];
      if constexpr (FINE) {
        const double s = gp(P.s64)[(long)ex * fny + ey];
        const int nny = fny + 1;
        double m[8];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const unsigned fb = fixbits(P.fix, (long)(ex + mg_lx(a)) * nny + (ey + mg_ly(a)));
          m[2 * a] = (fb & 1u) ? 0.0 : 1.0;
          m[2 * a + 1] = (fb & 2u) ? 0.0 : 1.0;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
          for (int j = i; j < 8; ++j) kf[sym8(i, j)] = s * P.Ke64[i * 8 + j] * m[i] * m[j];
      } else {
        const GLB T* K = gp((const T*)P.mg[l].K) + ((long)ex * fny + ey) * 36;
#pragma unroll
        for (int t = 0; t < 36; ++t) kf[t] = (double)K[t];
      }
      // kc += P_c^T kf P_c, one coarse column (kp, dp) at a time
#pragma unroll
      for (int kp = 0; kp < 4; ++kp)
#pragma unroll
        for (int dp = 0; dp < 2; ++dp) {
          double t1[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int ap = 0; ap < 4; ++ap) {
              const double w = mg_w(cx, cy, ap, kp);
              if (w != 0.0) acc = fma(kf[sym8(i, 2 * ap + dp)], w, acc);
            }
            t1[i] = acc;
          }
          const int col = 2 * kp + dp;
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
              const int row = 2 * k + d;
              if (row > col) continue;
              double acc = kc[sym8(row, col)];
#pragma unroll
              for (int a = 0; a < 4; ++a) {
                const double w = mg_w(cx, cy, a, k);
                if (w != 0.0) acc = fma(w, t1[2 * a + d], acc);
              }
              kc[sym8(row, col)] = acc;
            }
        }
    }
  GLB T* out = gp((T*)P.mg[l + 1].K) + E * 36;
#pragma unroll
  for (int t = 0; t < 36; ++t) out[t] = (T)kc[t];
}

// 1/diag of the assembled level-l operator (l >= 1); 0 on decoupled DOFs.
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_diag(const ProbDev* tab, const int2* map, ProbDev single, int l) {
  int lb;
  PD& P = grid_prob<B>(tab, map, single, blockIdx.x, lb);
  const int nx = P.mg[l].nx, ny = P.mg[l].ny;
  const long n = (long)lb * TOPT_BLOCK + threadIdx.x;
  if (n >= (long)(nx + 1) * (ny + 1)) return;
  const int ix = (int)(n / (ny + 1)), iy = (int)(n - (long)ix * (ny + 1));
  const GLB T* K = gp((const T*)P.mg[l].K);
  double dx = 0.0, dy = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int ex = ix - 1 + a, ey = iy - 1 + b;
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
      const int me = mg_me(a, b);
      const GLB T* k = K + ((long)ex * ny + ey) * 36;
      dx += (double)k[sym8(2 * me, 2 * me)];
      dy += (double)k[sym8(2 * me + 1, 2 * me + 1)];
    }
  vec2<T> d;
  d.x = dx > 0.0 ? (T)(1.0 / dx) : T(0);
  d.y = dy > 0.0 ? (T)(1.0 / dy) : T(0);
  gp((vec2<T>*)P.mg[l].dinv)[n] = d;
}

// Inverse of the coarsest operator (one workgroup, in LDS), formed per solve.
// The nodes are numbered along the longer side of the grid (permutation
// pos), which makes the matrix banded with half-bandwidth w = 2 (m + 1) + 1
// DOFs, m = nodes across the shorter side (<= 8 when N <= 128, so w <= 19).
// Steps: assemble A (decoupled DOFs get a unit diagonal); banded Cholesky
// A = L L^T in place (diagonal kept apart); copy the band of L to a compact
// array; solve L L^T x = e_c for every column c (one thread per column, the
// columns overwrite A); store A^-1 in the natural DOF order.
// Modified pivots: a pivot a_kk becomes max(|a_kk|, tau d_k), d_k the
// assembled diagonal. With FP32 level matrices the rounding of the Galerkin
// chain (relative 6e-8 per stored entry) exceeds the energy of near-null
// modes (a support held only through void elements has pivot ratios
// a_kk / d_k near 2e-8), so such pivots come out as noise of either sign, at
// about 1e-3 of d_k; tau = 1e-3 keeps the coarse solve symmetric positive
// definite and bounded there, as PCG needs. FP64 levels keep those pivots
// (tau = 1e-12); genuine pivots are far above tau in both cases.
#define TOPT_MG_SETUP_BLOCK 128   // >= TOPT_MG_DENSE: one thread per column
#define TOPT_MG_BAND 19           // max half-bandwidth (DOFs) for TOPT_MG_DENSE DOFs
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_MG_SETUP_BLOCK) void k_mg_coarse_setup(const ProbDev* tab, const int2* map,
                                                                         ProbDev single) {
  int lb;
  PD& P = grid_prob<B>(tab, map, single, blockIdx.x, lb);
  const int L = P.mg_nlev - 1;
  const int nx = P.mg[L].nx, ny = P.mg[L].ny, nnx = nx + 1, nny = ny + 1;
  const int N = P.mg_ncd;   // 2 nnx nny <= TOPT_MG_DENSE
  const bool tr = nny > nnx;
  const int m = tr ? nnx : nny;
  const int w = min(2 * (m + 1) + 1, N - 1);
  __shared__ double A[TOPT_MG_DENSE * TOPT_MG_DENSE];
  __shared__ double Lb[TOPT_MG_DENSE * (TOPT_MG_BAND + 1)];
  __shared__ double Ld[TOPT_MG_DENSE];
  __shared__ double D0[TOPT_MG_DENSE];
  const double tau = sizeof(T) == 4 ? 1e-3 : 1e-12;
  // permuted node p -> grid (ix, iy)
  auto node_of = [&](int p, int& ix, int& iy) {
    if (tr) { iy = p / nnx; ix = p - iy * nnx; }
    else { ix = p / nny; iy = p - ix * nny; }
  };
  const GLB T* K = gp((const T*)P.mg[L].K);
  for (int idx = threadIdx.x; idx < N * N; idx += TOPT_MG_SETUP_BLOCK) {
    const int i = idx / N, j = idx - (idx / N) * N;
    const int di = i & 1, dj = j & 1;
    int ix, iy, jx, jy;
    node_of(i >> 1, ix, iy);
    node_of(j >> 1, jx, jy);
    double v = 0.0;
    if (abs(ix - jx) <= 1 && abs(iy - jy) <= 1) {
      const int ex0 = max(ix, jx) - 1, ex1 = min(ix, jx);
      const int ey0 = max(iy, jy) - 1, ey1 = min(iy, jy);
      for (int ex = max(ex0, 0); ex <= min(ex1, nx - 1); ++ex)
        for (int ey = max(ey0, 0); ey <= min(ey1, ny - 1); ++ey) {
          const int li = (ix - ex) == 1 ? ((iy - ey) == 1 ? 2 : 1) : ((iy - ey) == 1 ? 3 : 0);
          const int lj = (jx - ex) == 1 ? ((jy - ey) == 1 ? 2 : 1) : ((jy - ey) == 1 ? 3 : 0);
          v += (double)K[((long)ex * ny + ey) * 36 + sym8(2 * li + di, 2 * lj + dj)];
        }
    }
    A[idx] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += TOPT_MG_SETUP_BLOCK) {
    if (!(A[i * N + i] > 0.0)) A[i * N + i] = 1.0;
    D0[i] = A[i * N + i];
  }
  __syncthreads();
  // banded Cholesky: column k scaled by 1 / l_kk, then the trailing band updated
  for (int k = 0; k < N; ++k) {
    const int kmax = min(N - 1, k + w);
    const double lkk = sqrt(fmax(fabs(A[k * N + k]), tau * D0[k]));
    if (threadIdx.x == 0) Ld[k] = lkk;
    for (int i = k + 1 + threadIdx.x; i <= kmax; i += TOPT_MG_SETUP_BLOCK) A[i * N + k] /= lkk;
    __syncthreads();
    const int nr = kmax - k;   // rows below the diagonal in the band
    for (int t = threadIdx.x; t < nr * (nr + 1) / 2; t += TOPT_MG_SETUP_BLOCK) {
      // t -> (i, j), k < j <= i <= kmax
      int ri = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
      while ((ri + 1) * (ri + 2) / 2 <= t) ++ri;
      while (ri * (ri + 1) / 2 > t) --ri;
      const int rj = t - ri * (ri + 1) / 2;
      const int i = k + 1 + ri, j = k + 1 + rj;
      A[i * N + j] = fma(-A[i * N + k], A[j * N + k], A[i * N + j]);
    }
    __syncthreads();
  }
  // compact band of L: Lb[i][d] = L(i, i - d), d = 1..w
  for (int idx = threadIdx.x; idx < N * w; idx += TOPT_MG_SETUP_BLOCK) {
    const int i = idx / w, d = idx - (idx / w) * w + 1;
    Lb[i * (TOPT_MG_BAND + 1) + d] = i - d >= 0 ? A[i * N + (i - d)] : 0.0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < N * N; idx += TOPT_MG_SETUP_BLOCK)
    A[idx] = (idx / N) == (idx - (idx / N) * N) ? 1.0 : 0.0;
  __syncthreads();
  // column c of the inverse: forward L y = e_c, then backward L^T x = y
  const int c = threadIdx.x;
  if (c < N) {
    for (int j = c; j < N; ++j) {
      const double yj = A[j * N + c] / Ld[j];
      A[j * N + c] = yj;
      for (int i = j + 1; i <= min(N - 1, j + w); ++i)
        A[i * N + c] = fma(-Lb[i * (TOPT_MG_BAND + 1) + (i - j)], yj, A[i * N + c]);
    }
    for (int j = N - 1; j >= 0; --j) {
      const double xj = A[j * N + c] / Ld[j];
      A[j * N + c] = xj;
      for (int i = max(0, j - w); i < j; ++i)
        A[i * N + c] = fma(-Lb[j * (TOPT_MG_BAND + 1) + (j - i)], xj, A[i * N + c]);
    }
  }
  __syncthreads();
  // back to the natural DOF order (node ix * nny + iy)
  for (int idx = threadIdx.x; idx < N * N; idx += TOPT_MG_SETUP_BLOCK) {
    const int i = idx / N, j = idx - (idx / N) * N;
    int ix, iy, jx, jy;
    node_of(i >> 1, ix, iy);
    node_of(j >> 1, jx, jy);
    const int ni = 2 * (ix * nny + iy) + (i & 1), nj = 2 * (jx * nny + jy) + (j & 1);
    gp(P.mg_inv)[(long)ni * N + nj] = A[idx];
  }
}

// ---------------------------------------------------------------- V-cycle
// The V-cycle kernels run between PCG slots B and C; cg[par] is the state
// slot B wrote. Once the solve has stopped they return at once (the host
// enqueues whole chunks of iterations).
__device__ __forceinline__ bool mg_stopped(PD& P, int par) {
  // the level-0 sweeps use the PCG tile map, which may hold problems of
  // the plan with another three-slot preconditioner
  return P.precond != PC_MG || cp((const CgState*)P.cg)[par].status != ST_RUN;
}
// The V-cycle kernels cover a batch through (problem, local block) maps: one
// map per level and kind, holding only the problems that take that step.
__device__ __forceinline__ PD& mg_entry(const ProbDev* tab, const int2* map, int b, int& lb) {
  const int2 e = map[b];
  lb = e.y;
  return *(PD*)(tab + e.x);
}
// Level l's corrected iterate after its up sweep: the coarsest level's in x,
// the others' in r.
__device__ __forceinline__ const void* mg_xup(PD& P, int l) {
  return l == P.mg_nlev - 1 ? P.mg[l].x : P.mg[l].r;
}
// The per-node sweep bodies take their node-vector pointers as template
// parameters: global memory in the per-level kernels, LDS in the tail.

// Restriction b_{l+1} = P^T r_l at coarse node N (level 0: fixed DOFs masked).
template <typename T, typename VR, typename VB>
__device__ __forceinline__ void mg_restrict_p(PD& P, int l, long N, VR r, VB bout) {
  const int fnx = P.mg[l].nx, fny = P.mg[l].ny;
  const int cny = P.mg[l + 1].ny;
  const int I = (int)(N / (cny + 1)), J = (int)(N - (long)I * (cny + 1));
  double bx = 0.0, by = 0.0;
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj) {
      const int i = 2 * I + di, j = 2 * J + dj;
      if (i < 0 || i > fnx || j < 0 || j > fny) continue;
      const double w = (di == 0 ? 1.0 : 0.5) * (dj == 0 ? 1.0 : 0.5);
      const long m = (long)i * (fny + 1) + j;
      const vec2<T> v = r[m];
      double vx = (double)v.x, vy = (double)v.y;
      if (l == 0) {
        const unsigned fb = fixbits(P.fix, m);
        if (fb & 1u) vx = 0.0;
        if (fb & 2u) vy = 0.0;
      }
      bx = fma(w, vx, bx);
      by = fma(w, vy, by);
    }
  vec2<T> o;
  o.x = (T)bx;
  o.y = (T)by;
  bout[N] = o;
}

template <typename T>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_restrict(const ProbDev* tab, const int2* map, int l, int par) {
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  if (mg_stopped(P, par)) return;
  const long N = (long)lb * TOPT_BLOCK + threadIdx.x;
  if (N >= (long)(P.mg[l + 1].nx + 1) * (P.mg[l + 1].ny + 1)) return;
  mg_restrict_p<T>(P, l, N, gp((const vec2<T>*)P.mg[l].r), gp((vec2<T>*)P.mg[l + 1].b));
}

// Interpolated coarse correction (P x_c) at fine node (i, j); xc holds the
// coarse level's iterate, cny its element rows.
template <typename T, typename VC>
__device__ __forceinline__ d2v mg_prolong(VC xc, int cny, int i, int j) {
  const int I = i >> 1, J = j >> 1;
  const bool ox = i & 1, oy = j & 1;
  const long c = (long)I * (cny + 1) + J;
  const vec2<T> a = xc[c];
  d2v v = d2v{(double)a.x, (double)a.y};
  if (ox) {
    const vec2<T> b = xc[c + cny + 1];
    v = 0.5 * (v + d2v{(double)b.x, (double)b.y});
  }
  if (oy) {
    const vec2<T> e = xc[c + 1];
    d2v w = d2v{(double)e.x, (double)e.y};
    if (ox) {
      const vec2<T> f = xc[c + cny + 2];
      w = 0.5 * (w + d2v{(double)f.x, (double)f.y});
    }
    v = 0.5 * (v + w);
  }
  return v;
}

// Sum over node (ix, iy)'s elements of its two rows of K_E times x_E, where
// xv(dx, dy) gives x at the neighbour (ix + dx, iy + dy) (level >= 1).
template <typename T, typename XF>
__device__ __forceinline__ d2v mg_apply_rows(PD& P, int l, int ix, int iy, XF xv) {
  const int nx = P.mg[l].nx, ny = P.mg[l].ny;
  const GLB T* K = gp((const T*)P.mg[l].K);
  double ax = 0.0, ay = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int ex = ix - 1 + a, ey = iy - 1 + b;
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
      const int me = mg_me(a, b);
      const GLB T* k = K + ((long)ex * ny + ey) * 36;
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        // element corner kk is neighbour (a - 1 + lx, b - 1 + ly) of the node
        const d2v xv_ = xv(a - 1 + mg_lx(kk), b - 1 + mg_ly(kk));
        ax = fma((double)k[sym8(2 * me, 2 * kk)], xv_.x, ax);
        ax = fma((double)k[sym8(2 * me, 2 * kk + 1)], xv_.y, ax);
        ay = fma((double)k[sym8(2 * me + 1, 2 * kk)], xv_.x, ay);
        ay = fma((double)k[sym8(2 * me + 1, 2 * kk + 1)], xv_.y, ay);
      }
    }
  return d2v{ax, ay};
}

// Down sweep at node n of level l >= 1: x = omega D^-1 b (from zero),
// r = b - K x.
template <typename T, typename VB, typename VX>
__device__ __forceinline__ void mg_down_p(PD& P, int l, long n, VB b, VX x, VX r) {
  const int nx = P.mg[l].nx, ny = P.mg[l].ny, nny = ny + 1;
  const int ix = (int)(n / nny), iy = (int)(n - (long)ix * nny);
  const GLB vec2<T>* dv = gp((const vec2<T>*)P.mg[l].dinv);
  d2v xs[3][3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const long m = nbr_node(ix, iy, dx - 1, dy - 1, nx, ny);   // clamped: meets void only
      const vec2<T> bb = b[m], dd = dv[m];
      xs[dx][dy] = d2v{TOPT_MG_OMEGA * (double)dd.x * (double)bb.x, TOPT_MG_OMEGA * (double)dd.y * (double)bb.y};
    }
  const d2v ax = mg_apply_rows<T>(P, l, ix, iy, [&](int dx, int dy) { return xs[dx + 1][dy + 1]; });
  const vec2<T> bn = b[n];
  vec2<T> xo, ro;
  xo.x = (T)xs[1][1].x;
  xo.y = (T)xs[1][1].y;
  ro.x = (T)((double)bn.x - ax.x);
  ro.y = (T)((double)bn.y - ax.y);
  x[n] = xo;
  r[n] = ro;
}

template <typename T>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_down(const ProbDev* tab, const int2* map, int l, int par) {
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  if (mg_stopped(P, par)) return;
  const long n = (long)lb * TOPT_BLOCK + threadIdx.x;
  if (n >= (long)(P.mg[l].nx + 1) * (P.mg[l].ny + 1)) return;
  mg_down_p<T>(P, l, n, gp((const vec2<T>*)P.mg[l].b), gp((vec2<T>*)P.mg[l].x), gp((vec2<T>*)P.mg[l].r));
}

// Up sweep at node n of level l >= 1 (not the coarsest): x' = x + P x_{l+1},
// out = x' + omega D^-1 (b - K x').
template <typename T, typename VX, typename VC, typename VB, typename VO>
__device__ __forceinline__ void mg_up_p(PD& P, int l, long n, VX x, VC xc, VB b, VO out) {
  const int nx = P.mg[l].nx, ny = P.mg[l].ny, nny = ny + 1;
  const int cny = P.mg[l + 1].ny;
  const int ix = (int)(n / nny), iy = (int)(n - (long)ix * nny);
  d2v xs[3][3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      int jx = ix + dx - 1, jy = iy + dy - 1;
      jx = jx < 0 ? 0 : (jx > nx ? nx : jx);
      jy = jy < 0 ? 0 : (jy > ny ? ny : jy);
      const vec2<T> xv = x[(long)jx * nny + jy];
      xs[dx][dy] = d2v{(double)xv.x, (double)xv.y} + mg_prolong<T>(xc, cny, jx, jy);
    }
  const d2v ax = mg_apply_rows<T>(P, l, ix, iy, [&](int dx, int dy) { return xs[dx + 1][dy + 1]; });
  const vec2<T> bn = b[n];
  const vec2<T> dd = gp((const vec2<T>*)P.mg[l].dinv)[n];
  vec2<T> o;
  o.x = (T)(xs[1][1].x + TOPT_MG_OMEGA * (double)dd.x * ((double)bn.x - ax.x));
  o.y = (T)(xs[1][1].y + TOPT_MG_OMEGA * (double)dd.y * ((double)bn.y - ax.y));
  out[n] = o;
}

// The corrected iterate goes to r (free after the restriction).
template <typename T>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_up(const ProbDev* tab, const int2* map, int l, int par) {
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  if (mg_stopped(P, par)) return;
  const long n = (long)lb * TOPT_BLOCK + threadIdx.x;
  if (n >= (long)(P.mg[l].nx + 1) * (P.mg[l].ny + 1)) return;
  mg_up_p<T>(P, l, n, gp((const vec2<T>*)P.mg[l].x), gp((const vec2<T>*)mg_xup(P, l + 1)),
             gp((const vec2<T>*)P.mg[l].b), gp((vec2<T>*)P.mg[l].r));
}

// Levels lt..L (the small ones) in one workgroup: down sweeps and
// restrictions to the coarsest level, its dense solve, up sweeps back to lt.
// The levels' x, b and r live in LDS, so a phase hands its results to the
// next through LDS and a workgroup barrier; element matrices, 1/diag and
// the coarse inverse are read from global memory. The kernel starts with
// the restriction of level lt-1's residual (written by the sweep launched
// before it), and level lt's corrected iterate goes back to global memory
// (mg_xup) for the up sweep after it.
// One workgroup per problem.
#define TOPT_MG_TAIL 1024        // levels with at most this many nodes (one per thread:
#define TOPT_MG_TAIL_BLOCK 1024  // each phase is one round trip)
#define TOPT_MG_TAIL_SUM 2560    // nodes of all tail levels together (LDS: 3 vectors each)
template <typename T>
__global__ __launch_bounds__(TOPT_MG_TAIL_BLOCK) void k_mg_tail(const ProbDev* tab, const int2* map, int par) {
  __shared__ vec2<T> sx[TOPT_MG_TAIL_SUM], sb[TOPT_MG_TAIL_SUM], sr[TOPT_MG_TAIL_SUM];
  int lb;
  PD& P = mg_entry(tab, map, blockIdx.x, lb);
  if (mg_stopped(P, par)) return;
  const int lt = P.mg_lt;
  const int L = P.mg_nlev - 1;
  auto nodes = [&](int l) { return (long)(P.mg[l].nx + 1) * (P.mg[l].ny + 1); };
  int off[TOPT_MG_MAXLEV];
  {
    int o = 0;
    for (int l = lt; l <= L; ++l) {
      off[l] = o;
      o += (int)nodes(l);
    }
  }
  // b of level lt: restriction of level lt-1's residual
  for (long n = threadIdx.x; n < nodes(lt); n += TOPT_MG_TAIL_BLOCK)
    mg_restrict_p<T>(P, lt - 1, n, gp((const vec2<T>*)P.mg[lt - 1].r), (vec2<T>*)(sb + off[lt]));
  __syncthreads();
  for (int l = lt; l < L; ++l) {
    for (long n = threadIdx.x; n < nodes(l); n += TOPT_MG_TAIL_BLOCK)
      mg_down_p<T>(P, l, n, (const vec2<T>*)(sb + off[l]), (vec2<T>*)(sx + off[l]), (vec2<T>*)(sr + off[l]));
    __syncthreads();
    for (long n = threadIdx.x; n < nodes(l + 1); n += TOPT_MG_TAIL_BLOCK)
      mg_restrict_p<T>(P, l, n, (const vec2<T>*)(sr + off[l]), (vec2<T>*)(sb + off[l + 1]));
    __syncthreads();
  }
  {
    // x = A^-1 b; A is symmetric: read column `row` (consecutive threads,
    // consecutive addresses)
    const int N = P.mg_ncd;
    const T* b = (const T*)(sb + off[L]);
    T* x = (T*)(sx + off[L]);
    for (int row = threadIdx.x; row < N; row += TOPT_MG_TAIL_BLOCK) {
      const GLB double* A = gp(P.mg_inv) + row;
      double acc = 0.0;
      for (int j = 0; j < N; ++j) acc = fma(A[(long)j * N], (double)b[j], acc);
      x[row] = (T)acc;
    }
  }
  __syncthreads();
  for (int l = L - 1; l >= lt; --l) {
    const vec2<T>* xc = (const vec2<T>*)(l + 1 == L ? sx + off[L] : sr + off[l + 1]);
    for (long n = threadIdx.x; n < nodes(l); n += TOPT_MG_TAIL_BLOCK)
      mg_up_p<T>(P, l, n, (const vec2<T>*)(sx + off[l]), xc, (const vec2<T>*)(sb + off[l]),
                 (vec2<T>*)(sr + off[l]));
    __syncthreads();
  }
  {
    const vec2<T>* src = (const vec2<T>*)(lt == L ? sx + off[L] : sr + off[lt]);
    GLB vec2<T>* dst = gp((vec2<T>*)mg_xup(P, lt));
    for (long n = threadIdx.x; n < nodes(lt); n += TOPT_MG_TAIL_BLOCK) dst[n] = src[n];
  }
}

// Level 0 down sweep on the PCG tiles: x0 = omega D^-1 r, r1 = r - K(s) x0
// (fixed rows identity, so r1 = 0 there).
template <typename T, bool KST>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_down0(const ProbDev* tab, const int2* map, int par) {
  using Tl = Tile<T>;
  constexpr int NPT = Tl::NPT;
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  if (mg_stopped(P, par)) return;
  __shared__ d2v hal[Tl::HN];
  __shared__ double esc[Tl::EN];
  const TileOrg o = tile_org<T>(P, lb);
  const GLB vec2<T>* r = gp((const vec2<T>*)P.r);
  const GLB vec2<T>* dv = gp((const vec2<T>*)P.pre);
  TileShare<T> S;
  tile_share<T>(P, o, S);
  vec2<T> orr[NPT], od[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    orr[j] = r[S.ids[j].n];
    od[j] = dv[S.ids[j].n];
  }
  const vec2<T> rr = r[S.rn], rd = dv[S.rn];
  d2v hv[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j)
    hv[j] = d2v{TOPT_MG_OMEGA * (double)od[j].x * (double)orr[j].x, TOPT_MG_OMEGA * (double)od[j].y * (double)orr[j].y};
  tile_store<T>(S, hv, d2v{TOPT_MG_OMEGA * (double)rd.x * (double)rr.x, TOPT_MG_OMEGA * (double)rd.y * (double)rr.y},
                hal, esc);
  __syncthreads();
  double ax[NPT], ay[NPT];
  tile_apply<T, KST>(P, hal, esc, ax, ay);
  GLB vec2<T>* x0 = gp((vec2<T>*)P.mg[0].x);
  GLB vec2<T>* r1 = gp((vec2<T>*)P.mg[0].r);
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    if (!S.ids[j].act) continue;
    vec2<T> xo, ro;
    xo.x = (T)hv[j].x;
    xo.y = (T)hv[j].y;
    ro.x = (S.fb[j] & 1u) ? T(0) : (T)((double)orr[j].x - ax[j]);
    ro.y = (S.fb[j] & 2u) ? T(0) : (T)((double)orr[j].y - ay[j]);
    x0[S.ids[j].n] = xo;
    r1[S.ids[j].n] = ro;
  }
}

// Level 0 up sweep: x' = x0 + (masked) P x_1, z = x' + omega D^-1 (r - K x'),
// z = 0 on fixed DOFs.
template <typename T, bool KST>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_up0(const ProbDev* tab, const int2* map, int par) {
  using Tl = Tile<T>;
  constexpr int NPT = Tl::NPT;
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  if (mg_stopped(P, par)) return;
  __shared__ d2v hal[Tl::HN];
  __shared__ double esc[Tl::EN];
  const TileOrg o = tile_org<T>(P, lb);
  const GLB vec2<T>* x0 = gp((const vec2<T>*)P.mg[0].x);
  const GLB vec2<T>* xc = gp((const vec2<T>*)mg_xup(P, 1));
  const int cny = P.mg[1].ny;
  TileShare<T> S;
  tile_share<T>(P, o, S);
  d2v hv[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    const vec2<T> xv = x0[S.ids[j].n];
    d2v c = mg_prolong<T>(xc, cny, S.ids[j].ix, S.ids[j].iy);
    if (S.fb[j] & 1u) c.x = 0.0;
    if (S.fb[j] & 2u) c.y = 0.0;
    hv[j] = d2v{(double)xv.x, (double)xv.y} + c;
  }
  d2v rv;
  {
    const long m = S.rn;
    const int nny = P.nely + 1;
    const int jx = (int)(m / nny), jy = (int)(m - (long)jx * nny);
    const vec2<T> xv = x0[m];
    d2v c = mg_prolong<T>(xc, cny, jx, jy);
    const unsigned fb = fixbits(P.fix, m);
    if (fb & 1u) c.x = 0.0;
    if (fb & 2u) c.y = 0.0;
    rv = d2v{(double)xv.x, (double)xv.y} + c;
  }
  tile_store<T>(S, hv, rv, hal, esc);
  __syncthreads();
  double ax[NPT], ay[NPT];
  tile_apply<T, KST>(P, hal, esc, ax, ay);
  const GLB vec2<T>* r = gp((const vec2<T>*)P.r);
  const GLB vec2<T>* dv = gp((const vec2<T>*)P.pre);
  GLB vec2<T>* z = gp((vec2<T>*)P.z);
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    if (!S.ids[j].act) continue;
    const long n = S.ids[j].n;
    const vec2<T> rr = r[n], dd = dv[n];
    vec2<T> zo;
    zo.x = (S.fb[j] & 1u) ? T(0) : (T)(hv[j].x + TOPT_MG_OMEGA * (double)dd.x * ((double)rr.x - ax[j]));
    zo.y = (S.fb[j] & 2u) ? T(0) : (T)(hv[j].y + TOPT_MG_OMEGA * (double)dd.y * ((double)rr.y - ay[j]));
    z[n] = zo;
  }
}

}  // namespace topt

namespace topt {

// ---------------------------------------------------------------- fused slot
// Plans whose problems all use multigrid run the level-0 up sweep and slot C
// as one kernel (one launch less per CG iteration; arithmetic of the
// separate kernels). Slot B is not fused with the down sweep: that sweep
// needs r at the tile's halo ring, which neighbouring workgroups of the
// same launch update, so it would read old or new values depending on
// workgroup timing.

// Level-0 up sweep + slot C: z = x0 + P x_1 + omega D^-1 (r - K x'), z = 0
// on fixed DOFs, partial r.z.
template <typename T, bool KST>
__global__ __launch_bounds__(TOPT_BLOCK) void k_slot_c_mg(const ProbDev* tab, const int2* pmap, int q) {
  using Tl = Tile<T>;
  constexpr int NPT = Tl::NPT;
  __shared__ d2v hal[Tl::HN];
  __shared__ double esc[Tl::EN];
  __shared__ double sh[16];
  int lb;
  PD& P = mg_entry(tab, pmap, xcd_remap(blockIdx.x, gridDim.x), lb);
  const TileOrg o = tile_org<T>(P, lb);
  const GLB vec2<T>* x0 = gp((const vec2<T>*)P.mg[0].x);
  const GLB vec2<T>* xc = gp((const vec2<T>*)mg_xup(P, 1));
  const GLB vec2<T>* r = gp((const vec2<T>*)P.r);
  const GLB vec2<T>* dv = gp((const vec2<T>*)P.pre);
  const int cny = P.mg[1].ny;
  // state-independent loads first
  PrevParts pp;
  parts_fetch(P, q, pp);
  TileShare<T> S;
  tile_share<T>(P, o, S);
  d2v hv[NPT];
  vec2<T> orr[NPT], od[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    const vec2<T> xv = x0[S.ids[j].n];
    d2v c = mg_prolong<T>(xc, cny, S.ids[j].ix, S.ids[j].iy);
    if (S.fb[j] & 1u) c.x = 0.0;
    if (S.fb[j] & 2u) c.y = 0.0;
    hv[j] = d2v{(double)xv.x, (double)xv.y} + c;
    orr[j] = r[S.ids[j].n];
    od[j] = dv[S.ids[j].n];
  }
  d2v rv;
  {
    const long m = S.rn;
    const int nny = P.nely + 1;
    const int jx = (int)(m / nny), jy = (int)(m - (long)jx * nny);
    const vec2<T> xv = x0[m];
    d2v c = mg_prolong<T>(xc, cny, jx, jy);
    const unsigned fb = fixbits(P.fix, m);
    if (fb & 1u) c.x = 0.0;
    if (fb & 2u) c.y = 0.0;
    rv = d2v{(double)xv.x, (double)xv.y} + c;
  }
  bool upd;
  const CgState st = slot_begin<T, SK_C>(P, lb, q, pp, sh, &upd);
  if (st.status != ST_RUN) return;
  tile_store<T>(S, hv, rv, hal, esc);
  __syncthreads();
  double ax[NPT], ay[NPT];
  tile_apply<T, KST>(P, hal, esc, ax, ay);
  GLB vec2<T>* z = gp((vec2<T>*)P.z);
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    if (!S.ids[j].act) continue;
    vec2<T> zo;
    zo.x = (S.fb[j] & 1u) ? T(0) : (T)(hv[j].x + TOPT_MG_OMEGA * (double)od[j].x * ((double)orr[j].x - ax[j]));
    zo.y = (S.fb[j] & 2u) ? T(0) : (T)(hv[j].y + TOPT_MG_OMEGA * (double)od[j].y * ((double)orr[j].y - ay[j]));
    z[S.ids[j].n] = zo;
    v += (double)orr[j].x * (double)zo.x + (double)orr[j].y * (double)zo.y;
  }
  slot_end(P, lb, q, v, 0.0, sh);
}

}  // namespace topt
