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
// solve. Smoother: damped Jacobi (omega from Ke, below), one sweep before and one after
// the coarse correction, so M is symmetric positive definite and plain PCG
// applies. The iteration count depends little on the grid size (with
// omega 0.6, a 300x100 bridge after 20 OC iterations took 69 iterations
// against 1,738 with Jacobi; the 1200x400 bridge takes 15-60 per solve).
#pragma once
#include <type_traits>

#include "fem_kernels.h"

namespace topt {

#ifndef TOPT_MG_DENSE
#define TOPT_MG_DENSE 128      // max DOFs of the coarsest level (inverted in LDS, FP64: 128 KB)
#endif
// Jacobi damping, per problem (P.mg_omega). Every element matrix (s_e Ke,
// with or without fixed-DOF masking, and every Galerkin agglomerate of them)
// has lambda_max(D_e^-1 K_e) = mu, the value of Ke itself (2.889 for the
// unit-square Q4 with nu = 0.3, 2.0 at nu = 0, 3.41 at nu = 0.45), so
// lambda_max(D^-1 K) <= mu on every level whatever the densities, and a
// damped Jacobi step converges in the energy norm for omega < 2 / mu. The
// host sets omega = TOPT_MG_OMEGA_FRAC * 2 / mu (0.672 at nu = 0.3), which
// keeps the V-cycle SPD for any Poisson ratio. On the 1200x400 bridge,
// 200 OC iterations took 10,146 / 9,980 / 9,869 / 9,853 CG iterations at
// omega = 0.62 / 0.65 / 0.67 / 0.68; a fixed 0.65 would leave the guaranteed
// range from nu = 0.35 on.
#ifndef TOPT_MG_OMEGA_FRAC
#define TOPT_MG_OMEGA_FRAC 0.97
#endif

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
// Element matrices of a level with ny element rows are stored entry-major
// per element column: packed entry t (sym8) of element (ex, ey) at
// mg_kel(ex, ey, ny) + t * ny, so threads on neighbouring elements of a
// column read neighbouring words.
__host__ __device__ inline long mg_kel(int ex, int ey, int ny) { return (long)ex * 36 * ny + ey; }

// ---------------------------------------------------------------- setup
// Coarse element matrices of level l+1 from level l (FINE: l = 0, children
// s_e Ke with fixed DOFs masked). One thread per coarse element, FP64 sums.
template <typename T, bool FINE>
__device__ __forceinline__ void mg_galerkin_body(PD& P, int lb, int l) {
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
        const GLB T* K = gp((const T*)P.mg[l].K) + mg_kel(ex, ey, fny);
#pragma unroll
        for (int t = 0; t < 36; ++t) kf[t] = (double)K[(long)t * fny];
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
  GLB T* out = gp((T*)P.mg[l + 1].K) + mg_kel(I, J, cny);
#pragma unroll
  for (int t = 0; t < 36; ++t) out[(long)t * cny] = (T)kc[t];
}
template <typename T, bool FINE, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_galerkin(const ProbDev* tab, const int2* map, ProbDev single,
                                                           int l) {
  int lb;
  PD& P = grid_prob<B>(tab, map, single, blockIdx.x, lb);
  mg_galerkin_body<T, FINE>(P, lb, l);
}

// 1/diag of the assembled level-l operator (l >= 1); 0 on decoupled DOFs.
// On levels 1 .. L-1 also the node records (TOPT_MG_ST) the sweeps apply: the
// node's two rows of the assembled operator, each coefficient the FP64 sum of
// its element entries in ascending element order. The entry coupling nodes i and
// j sums the same (symmetric) element entries in the same order from either
// side, so the assembled operator is bitwise symmetric.
template <typename T>
__device__ __forceinline__ void mg_diag_body(PD& P, int lb, int l) {
  const int nx = P.mg[l].nx, ny = P.mg[l].ny;
  const long n = (long)lb * TOPT_BLOCK + threadIdx.x;
  if (n >= (long)(nx + 1) * (ny + 1)) return;
  const int ix = (int)(n / (ny + 1)), iy = (int)(n - (long)ix * (ny + 1));
  const GLB T* K = gp((const T*)P.mg[l].K);
  if (P.mg[l].st == nullptr) {
    double dx = 0.0, dy = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int ex = ix - 1 + a, ey = iy - 1 + b;
        if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
        const int me = mg_me(a, b);
        const GLB T* k = K + mg_kel(ex, ey, ny);
        dx += (double)k[(long)sym8(2 * me, 2 * me) * ny];
        dy += (double)k[(long)sym8(2 * me + 1, 2 * me + 1) * ny];
      }
    vec2<T> d;
    d.x = dx > 0.0 ? (T)(1.0 / dx) : T(0);
    d.y = dy > 0.0 ? (T)(1.0 / dy) : T(0);
    gp((vec2<T>*)P.mg[l].dinv)[n] = d;
    return;
  }
  double A[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) A[i] = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int ex = ix - 1 + a, ey = iy - 1 + b;
      if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
      const int me = mg_me(a, b);
      const GLB T* k = K + mg_kel(ex, ey, ny);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        // element corner kk is neighbour (a - 1 + lx, b - 1 + ly) of the node
        const int q = 3 * (a + mg_lx(kk)) + (b + mg_ly(kk));
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
          for (int c = 0; c < 2; ++c) A[18 * d + 2 * q + c] += (double)k[(long)sym8(2 * me + d, 2 * kk + c) * ny];
      }
    }
  // the diagonal entries are the sums k_mg_diag forms above, in the same order
  const double dx = A[2 * 4], dy = A[18 + 2 * 4 + 1];
  vec2<T> d;
  d.x = dx > 0.0 ? (T)(1.0 / dx) : T(0);
  d.y = dy > 0.0 ? (T)(1.0 / dy) : T(0);
  gp((vec2<T>*)P.mg[l].dinv)[n] = d;
  typedef T Q __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int W = 16 / (int)sizeof(T);
  T rec[TOPT_MG_ST];
#pragma unroll
  for (int i = 0; i < 36; ++i) rec[i] = (T)A[i];
  rec[36] = d.x;
  rec[37] = d.y;
  rec[38] = T(0);
  rec[39] = T(0);
  GLB Q* o = gp((Q*)P.mg[l].st) + n * (TOPT_MG_ST / W);
#pragma unroll
  for (int i = 0; i < TOPT_MG_ST / W; ++i) {
    Q v;
#pragma unroll
    for (int w = 0; w < W; ++w) v[w] = rec[i * W + w];
    o[i] = v;
  }
}
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_diag(const ProbDev* tab, const int2* map, ProbDev single, int l) {
  int lb;
  PD& P = grid_prob<B>(tab, map, single, blockIdx.x, lb);
  mg_diag_body<T>(P, lb, l);
}

// The Galerkin product l -> l+1 and the diagonal and node records of level l
// (l >= 1) in one launch: blocks [0, nga) form the former, the rest the
// latter. Both only read level l's element matrices, which the launch before
// wrote; they write disjoint arrays.
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_level(const ProbDev* tab, const int2* gmap, const int2* dmap,
                                                        ProbDev single, int l, int nga) {
  int lb;
  if ((int)blockIdx.x < nga) {
    PD& P = grid_prob<B>(tab, gmap, single, blockIdx.x, lb);
    mg_galerkin_body<T, false>(P, lb, l);
  } else {
    PD& P = grid_prob<B>(tab, dmap, single, (int)blockIdx.x - nga, lb);
    mg_diag_body<T>(P, lb, l);
  }
}

// Inverse of the coarsest operator (one workgroup, in LDS), formed per solve.
// The nodes are numbered along the longer side of the grid (permutation
// pos), which makes the matrix banded with half-bandwidth w = 2 (m + 1) + 1
// DOFs, m = nodes across the shorter side (<= 8 when N <= 128, so w <= 19).
// Steps: stage the level's element matrices in LDS (in the band array, free
// until the factorization ends); assemble the lower band of A (decoupled
// DOFs get a unit diagonal); banded Cholesky A = L L^T in place, one barrier
// per column (column k is scaled by 1 / l_kk only when the band of L is
// copied out, so a step reads nothing it writes); then every column c of
// A^-1 by row-oriented substitutions L y = e_c, L^T x = y (one thread per
// column, each step a dot product over the band whose reads are issued
// together); store A^-1 in the natural DOF order.
// Modified pivots: a pivot a_kk becomes max(|a_kk|, tau d_k), d_k the
// assembled diagonal. With FP32 level matrices the rounding of the Galerkin
// chain (relative 6e-8 per stored entry) exceeds the energy of near-null
// modes (a support held only through void elements has pivot ratios
// a_kk / d_k near 2e-8), so such pivots come out as noise of either sign, at
// about 1e-3 of d_k; tau = 1e-3 keeps the coarse solve symmetric positive
// definite and bounded there, as PCG needs. FP64 levels keep those pivots
// (tau = 1e-12); genuine pivots are far above tau in both cases.
#define TOPT_MG_SETUP_BLOCK 256   // >= TOPT_MG_DENSE: one thread per column (the band assembly uses all)
#define TOPT_MG_BAND 19           // max half-bandwidth (DOFs) for TOPT_MG_DENSE DOFs
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_MG_SETUP_BLOCK) void k_mg_coarse_setup(const ProbDev* tab, const int2* map,
                                                                         ProbDev single) {
  constexpr int BW = TOPT_MG_BAND, NB = TOPT_MG_SETUP_BLOCK;
  int lb;
  PD& P = grid_prob<B>(tab, map, single, blockIdx.x, lb);
  const int L = P.mg_nlev - 1;
  const int nx = P.mg[L].nx, ny = P.mg[L].ny, nnx = nx + 1, nny = ny + 1;
  const int N = P.mg_ncd;   // 2 nnx nny <= TOPT_MG_DENSE
  const bool tr = nny > nnx;
  const int m = tr ? nnx : nny;
  const int w = min(2 * (m + 1) + 1, N - 1);
  __shared__ double A[TOPT_MG_DENSE * TOPT_MG_DENSE];
  __shared__ double Lb[TOPT_MG_DENSE * (BW + 1)];   // first: the element matrices
  __shared__ double Ld[TOPT_MG_DENSE], D0[TOPT_MG_DENSE];
  __shared__ int px[TOPT_MG_DENSE / 2], py[TOPT_MG_DENSE / 2];
  __shared__ short2 tri[BW * (BW + 1) / 2];   // (ri, rj), rj <= ri: the trailing band of a column
  // the coarsest level has fewer than TOPT_MG_DENSE / 2 elements
  static_assert(36 * (TOPT_MG_DENSE / 2) <= TOPT_MG_DENSE * (TOPT_MG_BAND + 1), "coarsest element matrices fit the band array");
  const double tau = sizeof(T) == 4 ? 1e-3 : 1e-12;
  const int t = (int)threadIdx.x;
  // element matrices (<= 49 elements: <= 64 nodes) and node coordinates of
  // the permuted numbering
  {
    constexpr int S = (36 * 49 + NB - 1) / NB;   // loads per thread, issued before the stores
    const GLB T* K = gp((const T*)P.mg[L].K);
    const int nk = nx * ny * 36;
    T v[S];
#pragma unroll
    for (int k = 0; k < S; ++k) v[k] = K[min(t + k * NB, nk - 1)];
#pragma unroll
    for (int k = 0; k < S; ++k)
      if (t + k * NB < nk) Lb[t + k * NB] = (double)v[k];
    for (int p = t; p < nnx * nny; p += NB) {
      if (tr) { py[p] = p / nnx; px[p] = p - (p / nnx) * nnx; }
      else { px[p] = p / nny; py[p] = p - (p / nny) * nny; }
    }
    for (int u = t; u < BW * (BW + 1) / 2; u += NB) {
      int ri = 0;
      while ((ri + 1) * (ri + 2) / 2 <= u) ++ri;
      tri[u] = make_short2((short)ri, (short)(u - ri * (ri + 1) / 2));
    }
  }
  __syncthreads();
  // lower band of A: entries (i, i - d), d = 0..w
  for (int u = t; u < N * (w + 1); u += NB) {
    const int i = u / (w + 1), j = i - (u - (u / (w + 1)) * (w + 1));
    if (j < 0) continue;
    const int ix = px[i >> 1], iy = py[i >> 1], jx = px[j >> 1], jy = py[j >> 1];
    const int di = i & 1, dj = j & 1;
    double v = 0.0;
    if (abs(ix - jx) <= 1 && abs(iy - jy) <= 1) {
      for (int ex = max(max(ix, jx) - 1, 0); ex <= min(min(ix, jx), nx - 1); ++ex)
        for (int ey = max(max(iy, jy) - 1, 0); ey <= min(min(iy, jy), ny - 1); ++ey) {
          const int li = (ix - ex) == 1 ? ((iy - ey) == 1 ? 2 : 1) : ((iy - ey) == 1 ? 3 : 0);
          const int lj = (jx - ex) == 1 ? ((jy - ey) == 1 ? 2 : 1) : ((jy - ey) == 1 ? 3 : 0);
          v += Lb[mg_kel(ex, ey, ny) + (long)sym8(2 * li + di, 2 * lj + dj) * ny];
        }
    }
    if (i == j && !(v > 0.0)) v = 1.0;
    A[i * N + j] = v;
    if (i == j) D0[i] = v;
  }
  __syncthreads();
  // banded Cholesky, column k: pivot, then the trailing band (k < j <= i <=
  // k + w) loses a_ik a_jk / a_kk; column k itself is left unscaled. (One
  // wave without block barriers, its loads batched, took 39-45 us here
  // against 31-34 us: the barrier is not what bounds a column.)
  for (int k = 0; k < N; ++k) {
    const double akk = fmax(fabs(A[k * N + k]), tau * D0[k]);
    const double inv = 1.0 / akk;
    if (t == 0) Ld[k] = sqrt(akk);
    const int nr = min(N - 1, k + w) - k;   // rows below the diagonal in the band
    for (int u = t; u < nr * (nr + 1) / 2; u += NB) {
      const short2 q = tri[u];
      const int i = k + 1 + q.x, j = k + 1 + q.y;
      A[i * N + j] = fma(-A[i * N + k] * inv, A[j * N + k], A[i * N + j]);
    }
    __syncthreads();
  }
  // band of L (Lb[i][d] = L(i, i - d), 0 for d > w or i - d < 0) and 1 / l_jj
  // (in D0)
  for (int u = t; u < N * BW; u += NB) {
    const int i = u / BW, d = u - (u / BW) * BW + 1;
    Lb[i * (BW + 1) + d] = d <= w && i - d >= 0 ? A[i * N + (i - d)] / Ld[i - d] : 0.0;
  }
  for (int i = t; i < N; i += NB) D0[i] = 1.0 / Ld[i];
  __syncthreads();
  // column c of the inverse: y_j = (e_c - sum_d L(j, j - d) y_{j-d}) / l_jj,
  // then x_j = (y_j - sum_d L(j + d, j) x_{j+d}) / l_jj; every lane runs all
  // N steps (y_j = 0 for j < c comes out of the recurrence). The last BW
  // values of a recurrence stay in registers: a ring of BW + 1 slots,
  // indexed statically by unrolling the steps by BW + 1, so a step reads
  // only the band of L (the same address for all lanes); the terms of
  // values older than the previous one are summed first, and the previous
  // value enters last, so consecutive steps overlap.
  const int c = t;
  if (c < N) {
    constexpr int R = BW + 1;
    double ring[R];
    // forward step j (ring slot r = j mod R)
    auto fwd = [&](int r, int j) {
      const double* lj = Lb + j * (BW + 1);
      double a0 = j == c ? 1.0 : 0.0, a1 = 0.0;
#pragma unroll
      for (int d = BW; d >= 2; --d) {
        // L(j, j - d) is 0 for d > w and for j - d < 0 (ring slot still 0)
        if (d & 1) a1 = fma(-lj[d], ring[(r - d + R) % R], a1);
        else a0 = fma(-lj[d], ring[(r - d + R) % R], a0);
      }
      const double y = fma(-lj[1], ring[(r - 1 + R) % R], a0 + a1) * D0[j];
      ring[r] = y;
      A[j * N + c] = y;
    };
    // backward step k, row j = N - 1 - k (x_{j+d} sits d steps back)
    auto bwd = [&](int r, int k) {
      const int j = N - 1 - k;
      double a0 = A[j * N + c], a1 = 0.0;
#pragma unroll
      for (int d = BW; d >= 2; --d) {
        // L(j + d, j): 0 for d > w; rows past N - 1 meet a ring slot still 0
        const int jj = j + d < N ? j + d : N - 1;
        if (d & 1) a1 = fma(-Lb[jj * (BW + 1) + d], ring[(r - d + R) % R], a1);
        else a0 = fma(-Lb[jj * (BW + 1) + d], ring[(r - d + R) % R], a0);
      }
      const int j1 = j + 1 < N ? j + 1 : N - 1;
      const double x = fma(-Lb[j1 * (BW + 1) + 1], ring[(r - 1 + R) % R], a0 + a1) * D0[j];
      ring[r] = x;
      A[j * N + c] = x;
    };
    // whole blocks of R steps carry no bounds test, so a block's reads can
    // be issued ahead of its recurrence; then the partial block
#pragma unroll
    for (int q = 0; q < R; ++q) ring[q] = 0.0;
    int j0 = 0;
    for (; j0 + R <= N; j0 += R) {
#pragma unroll
      for (int r = 0; r < R; ++r) fwd(r, j0 + r);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (j0 + r < N) fwd(r, j0 + r);
#pragma unroll
    for (int q = 0; q < R; ++q) ring[q] = 0.0;
    int k0 = 0;
    for (; k0 + R <= N; k0 += R) {
#pragma unroll
      for (int r = 0; r < R; ++r) bwd(r, k0 + r);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (k0 + r < N) bwd(r, k0 + r);
  }
  __syncthreads();
  // natural DOF order (node ix * nny + iy): row i by all threads together
  if (c < N) {
    const int nc = 2 * (px[c >> 1] * nny + py[c >> 1]) + (c & 1);
    for (int i = 0; i < N; ++i) {
      const int ni = 2 * (px[i >> 1] * nny + py[i >> 1]) + (i & 1);
      gp((T*)P.mg_inv)[(long)ni * N + nc] = (T)A[i * N + c];
    }
  }
}

// ---------------------------------------------------------------- V-cycle
// The V-cycle kernels run between PCG slots B and C; cg[par] is the state
// slot B wrote. Once the solve has stopped they return at once (the host
// enqueues whole chunks of iterations).
__device__ __forceinline__ bool mg_stopped(PD& P, int par) {
  // the level-0 sweeps use the PCG tile map, which may hold problems of
  // the plan with another three-slot preconditioner. A slot B that confirmed
  // a stop on the true residual (B_REF with a pending status) is followed by
  // the slot C that ends the solve: its V-cycle would go unused.
  // it, status, op and pending come in one 16-byte scalar load and combine
  // without branches (short-circuit tests would wait for each in turn).
  typedef int i4 __attribute__((ext_vector_type(4)));
  static_assert(offsetof(CgState, it) % 16 == 0 && offsetof(CgState, status) == offsetof(CgState, it) + 4 &&
                    offsetof(CgState, op) == offsetof(CgState, it) + 8 &&
                    offsetof(CgState, pending) == offsetof(CgState, it) + 12 && sizeof(CgState) % 16 == 0,
                "it, status, op, pending: one aligned 16-byte word");
  const i4 w = *(const CST i4*)&cp((const CgState*)P.cg)[par].it;
  return (int)(P.precond != PC_MG) | (int)(w.y != ST_RUN) | ((int)(w.z == OP_B_REF) & (int)(w.w != 0));
}
// The same test with the state slot passed as a kernel argument (one-problem
// plans, DIRECT): its load leaves with the descriptor's instead of after it.
template <bool DIRECT>
__device__ __forceinline__ bool mg_stopped(PD& P, int par, const CgState* cgd) {
  if constexpr (!DIRECT) return mg_stopped(P, par);
  typedef int i4 __attribute__((ext_vector_type(4)));
  const i4 w = *(const CST i4*)&cp(cgd)->it;
  return (int)(P.precond != PC_MG) | (int)(w.y != ST_RUN) | ((int)(w.z == OP_B_REF) & (int)(w.w != 0));
}
// The V-cycle kernels cover a batch through (problem, local block) maps: one
// map per level and kind, holding only the problems that take that step (a
// null map: the identity map of a one-problem plan).
__device__ __forceinline__ PD& mg_entry(const ProbDev* tab, const int2* map, int b, int& lb) {
  if (map == nullptr) {
    lb = b;
    return *(PD*)tab;
  }
  const int2 e = map[b];
  lb = e.y;
  return *(PD*)(tab + e.x);
}
// Level l's corrected iterate after its up sweep: the coarsest level's in x,
// the others' in r.
__device__ __forceinline__ const void* mg_xup(PD& P, int l) {
  return l == P.mg_nlev - 1 ? P.mg[l].x : P.mg[l].r;
}
// The per-node restriction and prolongation below also serve the tail
// kernel, whose node vectors are in LDS.

// Restriction b_{l+1} = P^T r_l at coarse node N (tail levels, l >= 1).
template <typename T, typename VR>
__device__ __forceinline__ vec2<T> mg_restrict_p(PD& P, int l, long N, VR r) {
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
      const vec2<T> v = r[(long)i * (fny + 1) + j];
      bx = fma(w, (double)v.x, bx);
      by = fma(w, (double)v.y, by);
    }
  vec2<T> o;
  o.x = (T)bx;
  o.y = (T)by;
  return o;
}

// Interpolated coarse correction (P x_c) at fine node (i, j); xc holds the
// coarse level's iterate, cny its element rows. Branch-free (odd and even
// nodes share a wavefront): the four reads are issued together, and on an
// even row or column the second of a pair repeats the first, so the pair's
// average is that value exactly and the result is bitwise the bilinear
// interpolation of the 1, 2 or 4 coarse values.
template <typename T, typename VC>
__device__ __forceinline__ d2v mg_prolong(VC xc, int cny, int i, int j) {
  const long c = (long)(i >> 1) * (cny + 1) + (j >> 1);
  const long dx = (i & 1) ? cny + 1 : 0, dy = j & 1;
  const vec2<T> a = xc[c], b = xc[c + dx], e = xc[c + dy], f = xc[c + dx + dy];
  const d2v v = 0.5 * (d2v{(double)a.x, (double)a.y} + d2v{(double)b.x, (double)b.y});
  const d2v w = 0.5 * (d2v{(double)e.x, (double)e.y} + d2v{(double)f.x, (double)f.y});
  return 0.5 * (v + w);
}

// Pre-smoothed iterate x = omega D^-1 b (from zero) at a node with 1/diag dd.
template <typename T>
__device__ __forceinline__ vec2<T> mg_jac(vec2<T> dd, vec2<T> bb, double om) {
  vec2<T> o;
  o.x = (T)(om * (double)dd.x * (double)bb.x);
  o.y = (T)(om * (double)dd.y * (double)bb.y);
  return o;
}

// ------------------------------------------------- down sweep + restriction
// The pre-smoothing step starts from zero, so level l's smoothed iterate
// x_l = omega D^-1 b_l is a pointwise function of b_l, and one kernel can
// take b_l to b_{l+1} = P^T (b_l - K_l x_l) reading only data written before
// its launch. A workgroup covers a tile of DR_CX x DR_CY coarse nodes: it
// forms x_l on the fine patch around them in LDS, r_l at the
// (2 DR_CX + 1) x (2 DR_CY + 1) fine nodes their restrictions gather (one
// per thread; the element rows each thread needs are fetched before the
// barrier), then the restrictions from LDS (level 0: fixed DOFs masked).
// Fine nodes on a tile's border are formed by both neighbouring tiles (245
// residuals per tile against its 4 x 48 = 192 fine nodes). x_l is not
// stored: the up sweep forms it again from b_l.
#ifndef TOPT_MG_DR_CX
#define TOPT_MG_DR_CX 2    // 2 x 24 coarse nodes: long runs along the memory order
#define TOPT_MG_DR_CY 24   // (7 x 8, 3 x 16 and 1 x 31 measured slower)
#endif
struct DrTile {
  static constexpr int CX = TOPT_MG_DR_CX, CY = TOPT_MG_DR_CY;   // coarse nodes
  static constexpr int RX = 2 * CX + 1, RY = 2 * CY + 1;         // fine residuals (5 x 49)
  static constexpr int XX = RX + 2, XY = RY + 2;                 // fine iterates (7 x 51)
};
static_assert(DrTile::RX * DrTile::RY <= TOPT_BLOCK, "one fine residual per thread");
// Workgroups of the fused sweep for a coarse level of cnx x cny elements.
__host__ __device__ inline int mg_dr_tiles(int cnx, int cny) {
  return ((cnx + DrTile::CX) / DrTile::CX) * ((cny + DrTile::CY) / DrTile::CY);
}

// A node record of level l >= 1 (k_mg_diag): ten 16-byte loads.
template <typename T>
__device__ __forceinline__ void mg_rec_load(const void* st, long n, T (&rec)[TOPT_MG_ST]) {
  typedef T Q __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int W = 16 / (int)sizeof(T);
  const GLB Q* src = gp((const Q*)st) + n * (TOPT_MG_ST / W);
#pragma unroll
  for (int i = 0; i < TOPT_MG_ST / W; ++i) {
    const Q v = src[i];
#pragma unroll
    for (int w = 0; w < W; ++w) rec[i * W + w] = v[w];
  }
}
// The record's two rows applied to x at the node's 3 x 3 neighbourhood, x(q)
// the value at neighbour q = 3 (dx + 1) + (dy + 1); neighbours outside the
// grid have zero coefficients.
template <typename T, typename XQ>
__device__ __forceinline__ d2v mg_rec_apply(const T (&rec)[TOPT_MG_ST], XQ x) {
  double ax = 0.0, ay = 0.0;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const d2v v = x(q);
    ax = fma((double)rec[2 * q], v.x, ax);
    ax = fma((double)rec[2 * q + 1], v.y, ax);
    ay = fma((double)rec[18 + 2 * q], v.x, ay);
    ay = fma((double)rec[18 + 2 * q + 1], v.y, ay);
  }
  return d2v{ax, ay};
}
// Level l >= 1 rows from the node records.
template <typename T>
struct MgRec {
  static constexpr bool kRec = true;
  const void* st;
};
// The per-element row type of an element policy (none for records).
template <typename EK, bool = EK::kRec>
struct ElemsRows {
  using R = typename EK::R;
};
template <typename EK>
struct ElemsRows<EK, true> {
  struct R {};
};

// Element rows of level 0: s_e Ke (KST: Ke from its six magnitudes); the
// factor is the element scale, 0 outside the grid.
template <typename T, bool KST>
struct MgElems0 {
  static constexpr bool kRec = false;
  PD& P;
  const GLB T* s;
  struct R {
    double f;
  };
  __device__ __forceinline__ R fetch(int ex, int ey, int) const {
    const bool ok = ex >= 0 && ex < P.nelx && ey >= 0 && ey < P.nely;
    const int cx = ex < 0 ? 0 : (ex >= P.nelx ? P.nelx - 1 : ex), cy = ey < 0 ? 0 : (ey >= P.nely ? P.nely - 1 : ey);
    const double v = (double)s[(long)cx * P.nely + cy];
    return R{ok ? v : 0.0};
  }
  __device__ __forceinline__ double at(const R&, int me, int d, int c) const {
    const int i = 2 * me + d;
    if constexpr (KST) return KS[i][c] < 0 ? -P.kc[-KS[i][c]] : P.kc[KS[i][c]];
    else return P.Ke64[i * 8 + c];
  }
};

// One tile (workgroup lb) of the fused sweep from a fine level of fnx x fny
// elements (b, 1/diag dv) to the coarse level of cnx x cny elements (bout).
template <typename T, bool MASK, typename EK, typename VB, typename VD, typename VO>
__device__ __forceinline__ void mg_down_restrict_tile(PD& P, int lb, int fnx, int fny, int cnx, int cny, VB b, VD dv,
                                                      const EK& ek, VO bout) {
  using D = DrTile;
  __shared__ d2v xs[D::XX * D::XY];
  __shared__ d2v rs[D::RX * D::RY];
  const int nty = (cny + D::CY) / D::CY;
  const int tx = lb / nty, ty = lb - tx * nty;
  const int I0 = tx * D::CX, J0 = ty * D::CY;
  const int fnny = fny + 1;
  const int t = (int)threadIdx.x;
  // this thread's fine residual node (residual tile position (qx, qy))
  const bool hasr = t < D::RX * D::RY;
  const int qx = hasr ? t / D::RY : 0, qy = hasr ? t - (t / D::RY) * D::RY : 0;
  const int i = 2 * I0 - 1 + qx, j = 2 * J0 - 1 + qy;
  const bool inr = hasr && i >= 0 && i <= fnx && j >= 0 && j <= fny;
  const int ic = i < 0 ? 0 : (i > fnx ? fnx : i), jc = j < 0 ? 0 : (j > fny ? fny : j);
  const long mn = (long)ic * fnny + jc;
  const vec2<T> bn = b[mn];
  [[maybe_unused]] T rec[TOPT_MG_ST];
  [[maybe_unused]] typename ElemsRows<EK>::R er[2][2];
  if constexpr (EK::kRec) {
    mg_rec_load<T>(ek.st, mn, rec);
  } else {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) er[a][c] = ek.fetch(i - 1 + a, j - 1 + c, mg_me(a, c));
  }
  unsigned fb = 0u;
  if constexpr (MASK) fb = fixbits(P.fix, mn);
  // x on the fine patch (clamped into the grid: a clamped value meets only
  // elements outside it, whose factor is 0); all of a thread's patch loads
  // are issued before its first LDS store (one round trip, not one per
  // patch row of threads)
  constexpr int NPU = (D::XX * D::XY + TOPT_BLOCK - 1) / TOPT_BLOCK;
  vec2<T> pb[NPU], pd[NPU];
#pragma unroll
  for (int k = 0; k < NPU; ++k) {
    const int u0 = t + k * TOPT_BLOCK;
    const int u = u0 < D::XX * D::XY ? u0 : D::XX * D::XY - 1;
    const int px = u / D::XY, py = u - (u / D::XY) * D::XY;
    int jx = 2 * I0 - 2 + px, jy = 2 * J0 - 2 + py;
    jx = jx < 0 ? 0 : (jx > fnx ? fnx : jx);
    jy = jy < 0 ? 0 : (jy > fny ? fny : jy);
    const long m = (long)jx * fnny + jy;
    pb[k] = b[m];
    pd[k] = dv[m];
  }
#pragma unroll
  for (int k = 0; k < NPU; ++k) {
    const int u = t + k * TOPT_BLOCK;
    if (u < D::XX * D::XY)
      xs[u] = d2v{P.mg_omega * (double)pd[k].x * (double)pb[k].x, P.mg_omega * (double)pd[k].y * (double)pb[k].y};
  }
  __syncthreads();
  if (hasr) {
    double ax = 0.0, ay = 0.0;
    if constexpr (EK::kRec) {
      // neighbour q of the node sits at patch position (qx + 1 + dx, qy + 1 + dy)
      const d2v a2 = mg_rec_apply<T>(rec, [&](int q) { return xs[(qx + q / 3) * D::XY + qy + q % 3]; });
      ax = a2.x;
      ay = a2.y;
    } else {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int me = mg_me(a, c);
          double tx_ = 0.0, ty_ = 0.0;
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            // corner kk of element (i - 1 + a, j - 1 + c): patch position below
            const d2v v = xs[(qx + a + mg_lx(kk)) * D::XY + qy + c + mg_ly(kk)];
            tx_ = fma(ek.at(er[a][c], me, 0, 2 * kk), v.x, tx_);
            tx_ = fma(ek.at(er[a][c], me, 0, 2 * kk + 1), v.y, tx_);
            ty_ = fma(ek.at(er[a][c], me, 1, 2 * kk), v.x, ty_);
            ty_ = fma(ek.at(er[a][c], me, 1, 2 * kk + 1), v.y, ty_);
          }
          ax = fma(er[a][c].f, tx_, ax);
          ay = fma(er[a][c].f, ty_, ay);
        }
    }
    d2v r = d2v{(double)bn.x - ax, (double)bn.y - ay};
    if (!inr || (fb & 1u)) r.x = 0.0;
    if (!inr || (fb & 2u)) r.y = 0.0;
    rs[t] = r;
  }
  __syncthreads();
  if (t < D::CX * D::CY) {
    const int cx = t / D::CY, cy = t - (t / D::CY) * D::CY;
    const int I = I0 + cx, J = J0 + cy;
    if (I <= cnx && J <= cny) {
      double bx = 0.0, by = 0.0;
#pragma unroll
      for (int di = 0; di < 3; ++di)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          const double w = (di == 1 ? 1.0 : 0.5) * (dj == 1 ? 1.0 : 0.5);
          const d2v r = rs[(2 * cx + di) * D::RY + 2 * cy + dj];
          bx = fma(w, r.x, bx);
          by = fma(w, r.y, by);
        }
      vec2<T> o;
      o.x = (T)bx;
      o.y = (T)by;
      bout[(long)I * (cny + 1) + J] = o;
    }
  }
}

// Levels 1 .. lt-1: b_{l+1} from b_l.
template <typename T, bool DIRECT>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_down_restrict(const CgState* cgd, const ProbDev* tab, const int2* map,
                                                                 int l, int par) {
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  // the level's fields enter the test, so they are read with the state (one
  // scalar round trip; a level without buffers has nothing to sweep)
  const int fnx = P.mg[l].nx, fny = P.mg[l].ny, cnx = P.mg[l + 1].nx, cny = P.mg[l + 1].ny;
  const void *st = P.mg[l].st, *b = P.mg[l].b, *dv = P.mg[l].dinv, *bc = P.mg[l + 1].b;
  if ((int)mg_stopped<DIRECT>(P, par, cgd) | (int)(st == nullptr) | (int)(b == nullptr) | (int)(dv == nullptr) |
      (int)(bc == nullptr) | (int)((lb | fnx | fny | cnx | cny) < 0))
    return;
  const MgRec<T> ek{st};
  mg_down_restrict_tile<T, false>(P, lb, fnx, fny, cnx, cny, gp((const vec2<T>*)b), gp((const vec2<T>*)dv), ek,
                                  gp((vec2<T>*)bc));
}

// Two down sweeps in one launch: levels l and l+1 (the last two above the
// tail of a one-problem plan), b_{l+2} from b_l. A workgroup covers a tile of
// Dr2Tile::CX x CY level-(l+2) nodes. It forms b_{l+1} on the level-(l+1)
// patch its own sweep reads (XX2 x XY2 nodes) as k_mg_down_restrict forms it
// at level l (x_l on the level-l patch, the residuals from the node records,
// the restriction sums, rounded to T), stores the level-(l+1) nodes it owns
// (rows [2 I0, 2 I0 + 2 CX), columns [2 J0, 2 J0 + 2 CY) of that level: every
// node once over the tiles) for the up sweep, then sweeps level l+1 from the
// patch. Each value is formed by the same operations on the same operands as
// in the two separate sweeps, so the bits are theirs; patch positions
// outside level l+1 hold 0 where the separate sweep holds a clamped node's
// value, and both meet only zero coefficients there.
struct Dr2Tile {
  static constexpr int CX = 2, CY = 8;                       // level l+2 nodes
  static constexpr int RX2 = 2 * CX + 1, RY2 = 2 * CY + 1;   // level l+1 residuals (5 x 17)
  static constexpr int XX2 = RX2 + 2, XY2 = RY2 + 2;         // level l+1 patch (7 x 19)
  static constexpr int RX1 = 2 * XX2 + 1, RY1 = 2 * XY2 + 1; // level l residuals (15 x 39)
  static constexpr int XX1 = RX1 + 2, XY1 = RY1 + 2;         // level l patch (17 x 41)
};
static_assert(Dr2Tile::RX2 * Dr2Tile::RY2 <= TOPT_BLOCK && Dr2Tile::XX2 * Dr2Tile::XY2 <= TOPT_BLOCK,
              "one level-(l+1) residual and patch node per thread");
// Workgroups of the two-level sweep for a level l+2 of cnx x cny elements.
__host__ __device__ inline int mg_dr2_tiles(int cnx, int cny) {
  return ((cnx + Dr2Tile::CX) / Dr2Tile::CX) * ((cny + Dr2Tile::CY) / Dr2Tile::CY);
}
template <typename T, bool DIRECT>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_down2(const CgState* cgd, const ProbDev* tab, const int2* map,
                                                         int l, int par) {
  using D = Dr2Tile;
  constexpr int N1R = (D::RX1 * D::RY1 + TOPT_BLOCK - 1) / TOPT_BLOCK;   // level-l residuals per thread
  constexpr int N1X = (D::XX1 * D::XY1 + TOPT_BLOCK - 1) / TOPT_BLOCK;   // level-l patch nodes per thread
  __shared__ d2v xs1[D::XX1 * D::XY1];
  __shared__ d2v rs1[D::RX1 * D::RY1];
  __shared__ vec2<T> b2[D::XX2 * D::XY2];
  __shared__ d2v xs2[D::XX2 * D::XY2];
  __shared__ d2v rs2[D::RX2 * D::RY2];
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  const int n0x = P.mg[l].nx, n0y = P.mg[l].ny, n1x = P.mg[l + 1].nx, n1y = P.mg[l + 1].ny;
  const int n2x = P.mg[l + 2].nx, n2y = P.mg[l + 2].ny;
  const void *st0 = P.mg[l].st, *b0 = P.mg[l].b, *dv0 = P.mg[l].dinv;
  const void *st1 = P.mg[l + 1].st, *b1 = P.mg[l + 1].b, *dv1 = P.mg[l + 1].dinv, *bo = P.mg[l + 2].b;
  if ((int)mg_stopped<DIRECT>(P, par, cgd) | (int)(st0 == nullptr) | (int)(b0 == nullptr) | (int)(dv0 == nullptr) |
      (int)(st1 == nullptr) | (int)(b1 == nullptr) | (int)(dv1 == nullptr) | (int)(bo == nullptr) |
      (int)((lb | n0x | n0y | n1x | n1y | n2x | n2y) < 0))
    return;
  const GLB vec2<T>* bl = gp((const vec2<T>*)b0);
  const GLB vec2<T>* dl = gp((const vec2<T>*)dv0);
  const GLB vec2<T>* dm = gp((const vec2<T>*)dv1);
  const int nty = (n2y + D::CY) / D::CY;
  const int tx = lb / nty, ty = lb - tx * nty;
  const int I0 = tx * D::CX, J0 = ty * D::CY;                  // level l+2
  const int t = (int)threadIdx.x;
  const int n0ny = n0y + 1, n1ny = n1y + 1;
  // level l+1 residual node of this thread (its record goes first: it does
  // not depend on the level-l stage)
  const bool has2 = t < D::RX2 * D::RY2;
  const int q2x = has2 ? t / D::RY2 : 0, q2y = has2 ? t - (t / D::RY2) * D::RY2 : 0;
  const int i2 = 2 * I0 - 1 + q2x, j2 = 2 * J0 - 1 + q2y;
  const bool in2 = has2 && i2 >= 0 && i2 <= n1x && j2 >= 0 && j2 <= n1y;
  const long m2 = (long)(i2 < 0 ? 0 : (i2 > n1x ? n1x : i2)) * n1ny + (j2 < 0 ? 0 : (j2 > n1y ? n1y : j2));
  T rec2[TOPT_MG_ST];
  mg_rec_load<T>(st1, m2, rec2);
  // level-(l+1) patch node of this thread: its 1/diag
  const bool hasp = t < D::XX2 * D::XY2;
  const int ppx = hasp ? t / D::XY2 : 0, ppy = hasp ? t - (t / D::XY2) * D::XY2 : 0;
  const int kx = 2 * I0 - 2 + ppx, ky = 2 * J0 - 2 + ppy;      // level l+1
  const bool inp = hasp && kx >= 0 && kx <= n1x && ky >= 0 && ky <= n1y;
  const vec2<T> pd2 = dm[(long)(kx < 0 ? 0 : (kx > n1x ? n1x : kx)) * n1ny + (ky < 0 ? 0 : (ky > n1y ? n1y : ky))];
  // level-l residual nodes of this thread: records and b
  T rec1[N1R][TOPT_MG_ST];
  vec2<T> bn1[N1R];
  int q1x[N1R], q1y[N1R];
  bool in1[N1R];
#pragma unroll
  for (int k = 0; k < N1R; ++k) {
    const int u0 = t + k * TOPT_BLOCK;
    const int u = u0 < D::RX1 * D::RY1 ? u0 : D::RX1 * D::RY1 - 1;
    q1x[k] = u / D::RY1;
    q1y[k] = u - (u / D::RY1) * D::RY1;
    const int i = 4 * I0 - 5 + q1x[k], j = 4 * J0 - 5 + q1y[k];
    in1[k] = u0 < D::RX1 * D::RY1 && i >= 0 && i <= n0x && j >= 0 && j <= n0y;
    const long mn = (long)(i < 0 ? 0 : (i > n0x ? n0x : i)) * n0ny + (j < 0 ? 0 : (j > n0y ? n0y : j));
    bn1[k] = bl[mn];
    mg_rec_load<T>(st0, mn, rec1[k]);
  }
  // x_l on the level-l patch (clamped into the grid, as the separate sweep)
  vec2<T> pb[N1X], pd[N1X];
#pragma unroll
  for (int k = 0; k < N1X; ++k) {
    const int u0 = t + k * TOPT_BLOCK;
    const int u = u0 < D::XX1 * D::XY1 ? u0 : D::XX1 * D::XY1 - 1;
    const int px = u / D::XY1, py = u - (u / D::XY1) * D::XY1;
    int jx = 4 * I0 - 6 + px, jy = 4 * J0 - 6 + py;
    jx = jx < 0 ? 0 : (jx > n0x ? n0x : jx);
    jy = jy < 0 ? 0 : (jy > n0y ? n0y : jy);
    const long m = (long)jx * n0ny + jy;
    pb[k] = bl[m];
    pd[k] = dl[m];
  }
#pragma unroll
  for (int k = 0; k < N1X; ++k) {
    const int u = t + k * TOPT_BLOCK;
    if (u < D::XX1 * D::XY1)
      xs1[u] = d2v{P.mg_omega * (double)pd[k].x * (double)pb[k].x, P.mg_omega * (double)pd[k].y * (double)pb[k].y};
  }
  __syncthreads();
  // level-l residuals
#pragma unroll
  for (int k = 0; k < N1R; ++k) {
    const int u = t + k * TOPT_BLOCK;
    if (u < D::RX1 * D::RY1) {
      const int qx = q1x[k], qy = q1y[k];
      const d2v a2 = mg_rec_apply<T>(rec1[k], [&](int q) { return xs1[(qx + q / 3) * D::XY1 + qy + q % 3]; });
      d2v r = d2v{(double)bn1[k].x - a2.x, (double)bn1[k].y - a2.y};
      if (!in1[k]) r = d2v{0.0, 0.0};
      rs1[u] = r;
    }
  }
  __syncthreads();
  // b_{l+1} on the level-(l+1) patch (rounded to T), stored where owned
  if (hasp) {
    vec2<T> o;
    o.x = T(0);
    o.y = T(0);
    if (inp) {
      double bx = 0.0, by = 0.0;
#pragma unroll
      for (int di = 0; di < 3; ++di)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          const double w = (di == 1 ? 1.0 : 0.5) * (dj == 1 ? 1.0 : 0.5);
          const d2v r = rs1[(2 * ppx + di) * D::RY1 + 2 * ppy + dj];
          bx = fma(w, r.x, bx);
          by = fma(w, r.y, by);
        }
      o.x = (T)bx;
      o.y = (T)by;
      if (ppx >= 2 && ppx < 2 + 2 * D::CX && ppy >= 2 && ppy < 2 + 2 * D::CY)
        gp((vec2<T>*)b1)[(long)kx * n1ny + ky] = o;
    }
    b2[t] = o;
    xs2[t] = d2v{P.mg_omega * (double)pd2.x * (double)o.x, P.mg_omega * (double)pd2.y * (double)o.y};
  }
  __syncthreads();
  // level-(l+1) residuals
  if (has2) {
    const vec2<T> bn = b2[(q2x + 1) * D::XY2 + q2y + 1];
    const d2v a2 = mg_rec_apply<T>(rec2, [&](int q) { return xs2[(q2x + q / 3) * D::XY2 + q2y + q % 3]; });
    d2v r = d2v{(double)bn.x - a2.x, (double)bn.y - a2.y};
    if (!in2) r = d2v{0.0, 0.0};
    rs2[t] = r;
  }
  __syncthreads();
  if (t < D::CX * D::CY) {
    const int cx = t / D::CY, cy = t - (t / D::CY) * D::CY;
    const int I = I0 + cx, J = J0 + cy;
    if (I <= n2x && J <= n2y) {
      double bx = 0.0, by = 0.0;
#pragma unroll
      for (int di = 0; di < 3; ++di)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          const double w = (di == 1 ? 1.0 : 0.5) * (dj == 1 ? 1.0 : 0.5);
          const d2v r = rs2[(2 * cx + di) * D::RY2 + 2 * cy + dj];
          bx = fma(w, r.x, bx);
          by = fma(w, r.y, by);
        }
      vec2<T> o;
      o.x = (T)bx;
      o.y = (T)by;
      gp((vec2<T>*)bo)[(long)I * (n2y + 1) + J] = o;
    }
  }
}

// Level 0: b_1 from the PCG residual r (x0 = omega D^-1 r; K(s) has fixed
// rows identity, so r1 = 0 on fixed DOFs, and r, hence x0, is zero there).
template <typename T, bool KST, bool DIRECT>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_down_restrict0(const CgState* cgd, const ProbDev* tab,
                                                                  const int2* map, int par) {
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  // the fields enter the test, so they are read with the state
  const int fnx = P.nelx, fny = P.nely, cnx = P.mg[1].nx, cny = P.mg[1].ny;
  const void *r = P.r, *dv = P.pre, *sc = P.s, *bc = P.mg[1].b;
  if ((int)mg_stopped<DIRECT>(P, par, cgd) | (int)(r == nullptr) | (int)(dv == nullptr) | (int)(sc == nullptr) |
      (int)(bc == nullptr) | (int)((lb | fnx | fny | cnx | cny) < 0))
    return;
  const MgElems0<T, KST> ek{P, gp((const T*)sc)};
  mg_down_restrict_tile<T, true>(P, lb, fnx, fny, cnx, cny, gp((const vec2<T>*)r), gp((const vec2<T>*)dv), ek,
                                 gp((vec2<T>*)bc));
}

// Up sweep at levels 1 .. lt-1 on tiles of UpTile::TX x UpTile::TY nodes
// (one per thread): x' = omega D^-1 b + P x_{l+1} (the pre-smoothed iterate
// formed again from b) on the tile's patch in LDS, the element rows of each
// thread's node fetched before the barrier, then out = x' + omega D^-1
// (b - K x'). The corrected iterate goes to r (the fused down sweep stores
// no residual).
#ifndef TOPT_MG_UP_TX
#define TOPT_MG_UP_TX 4    // one node column per wavefront (8 x 32 and 16 x 16
#define TOPT_MG_UP_TY 64   // measured slower)
#endif
struct UpTile {
  static constexpr int TX = TOPT_MG_UP_TX, TY = TOPT_MG_UP_TY;   // nodes (TY along the memory order)
  static constexpr int PX = TX + 2, PY = TY + 2;     // patch with its ring
};
static_assert(UpTile::TX * UpTile::TY == TOPT_BLOCK, "one node per thread");
// Workgroups of the up sweep for a level of nx x ny elements.
__host__ __device__ inline int mg_up_tiles(int nx, int ny) {
  return ((nx + UpTile::TX) / UpTile::TX) * ((ny + UpTile::TY) / UpTile::TY);
}
template <typename T, bool DIRECT>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_up(const CgState* cgd, const ProbDev* tab, const int2* map, int l,
                                                      int par) {
  using U = UpTile;
  __shared__ d2v xs[U::PX * U::PY];
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  // the level's fields enter the test, so they are read with the state
  const int nx = P.mg[l].nx, ny = P.mg[l].ny, cny = P.mg[l + 1].ny;
  const void *st = P.mg[l].st, *bl = P.mg[l].b, *dl = P.mg[l].dinv, *rl = P.mg[l].r, *xup = mg_xup(P, l + 1);
  if ((int)mg_stopped<DIRECT>(P, par, cgd) | (int)(st == nullptr) | (int)(bl == nullptr) | (int)(dl == nullptr) |
      (int)(rl == nullptr) | (int)(xup == nullptr) | (int)((lb | nx | ny | cny) < 0))
    return;
  const int nny = ny + 1;
  const int nty = (ny + U::TY) / U::TY;
  const int tx = lb / nty, ty = lb - tx * nty;
  const int x0 = tx * U::TX, y0 = ty * U::TY;
  const int t = (int)threadIdx.x;
  const int ox = t / U::TY, oy = t - (t / U::TY) * U::TY;
  const int ix = x0 + ox, iy = y0 + oy;
  const bool act = ix <= nx && iy <= ny;
  const long n = (long)(ix <= nx ? ix : nx) * nny + (iy <= ny ? iy : ny);
  const GLB vec2<T>* b = gp((const vec2<T>*)bl);
  const GLB vec2<T>* dv = gp((const vec2<T>*)dl);
  const GLB vec2<T>* xc = gp((const vec2<T>*)xup);
  const vec2<T> bn = b[n];
  T rec[TOPT_MG_ST];   // the node's rows and 1/diag
  mg_rec_load<T>(st, n, rec);
  // x' on the patch (clamped into the grid: a clamped value meets only
  // neighbours outside it, whose coefficients are 0); all of a thread's patch
  // loads are issued before its first LDS store
  constexpr int NPU = (U::PX * U::PY + TOPT_BLOCK - 1) / TOPT_BLOCK;
  vec2<T> pb[NPU], pd[NPU];
  d2v pc[NPU];
#pragma unroll
  for (int k = 0; k < NPU; ++k) {
    const int u0 = t + k * TOPT_BLOCK;
    const int u = u0 < U::PX * U::PY ? u0 : U::PX * U::PY - 1;
    const int px = u / U::PY, py = u - (u / U::PY) * U::PY;
    int jx = x0 - 1 + px, jy = y0 - 1 + py;
    jx = jx < 0 ? 0 : (jx > nx ? nx : jx);
    jy = jy < 0 ? 0 : (jy > ny ? ny : jy);
    const long m = (long)jx * nny + jy;
    pb[k] = b[m];
    pd[k] = dv[m];
    pc[k] = mg_prolong<T>(xc, cny, jx, jy);
  }
#pragma unroll
  for (int k = 0; k < NPU; ++k) {
    const int u = t + k * TOPT_BLOCK;
    if (u < U::PX * U::PY)
      xs[u] = d2v{P.mg_omega * (double)pd[k].x * (double)pb[k].x, P.mg_omega * (double)pd[k].y * (double)pb[k].y} +
              pc[k];
  }
  __syncthreads();
  if (!act) return;
  // neighbour q of the node sits at patch position (ox + 1 + dx, oy + 1 + dy)
  const d2v ax = mg_rec_apply<T>(rec, [&](int q) { return xs[(ox + q / 3) * U::PY + oy + q % 3]; });
  const d2v xo = xs[(ox + 1) * U::PY + oy + 1];
  vec2<T> o;
  o.x = (T)(xo.x + P.mg_omega * (double)rec[36] * ((double)bn.x - ax.x));
  o.y = (T)(xo.y + P.mg_omega * (double)rec[37] * ((double)bn.y - ax.y));
  gp((vec2<T>*)rl)[n] = o;
}

// Levels lt..L (the small ones) in one workgroup: down sweeps and
// restrictions to the coarsest level, its dense solve, up sweeps back to lt.
// The levels' x, b and r live in LDS, so a phase hands its results to the
// next through LDS and a workgroup barrier. Every node of the sweep levels
// lt..L-1 has a slot (levels in order, nodes in order) in the LDS vectors,
// and slot g belongs to thread g mod TOPT_MG_TAIL_BLOCK: the thread applies
// the node's two rows of the assembled level operator from its record
// (written by k_mg_diag), 18 neighbour values per row. With FP32 vectors a
// thread's first record and its share of the coarse inverse come into
// registers on entry, all loads issued together, so the later phases read
// only LDS; with FP64 (no registers to spare) each phase reads them from
// global memory. The kernel starts from b_lt (written by the fused sweep
// launched before it), and level lt's corrected iterate goes back to global
// memory (mg_xup) for the up sweep after it. One workgroup per problem.
#ifndef TOPT_MG_TAIL
#define TOPT_MG_TAIL 1024        // levels with at most this many nodes
#endif
#define TOPT_MG_TAIL_BLOCK 1024
#ifndef TOPT_MG_TAIL_SUM
#define TOPT_MG_TAIL_SUM 2560    // nodes of all tail levels together (LDS: 3 vectors each)
#endif

// Slot g of the tail: its level, node, node position, the level's offset in
// the LDS vectors and its node records.
struct TailSlot {
  int l, n, ix, iy, nx, ny, off;
  const void* st;
};
__device__ __forceinline__ TailSlot tail_slot(PD& P, int lt, int L, int g) {
  TailSlot s{lt, g, 0, 0, P.mg[lt].nx, P.mg[lt].ny, 0, P.mg[lt].st};
  int acc = 0;
  for (int k = lt; k < L; ++k) {   // uniform loop over the levels, per-lane selects
    if (g >= acc) {
      s.l = k;
      s.n = g - acc;
      s.nx = P.mg[k].nx;
      s.ny = P.mg[k].ny;
      s.off = acc;
      s.st = P.mg[k].st;
    }
    acc += (P.mg[k].nx + 1) * (P.mg[k].ny + 1);
  }
  s.ix = s.n / (s.ny + 1);
  s.iy = s.n - s.ix * (s.ny + 1);
  return s;
}
template <typename T>
__device__ __forceinline__ void tail_rec_load(const TailSlot& s, T (&rec)[TOPT_MG_ST]) {
  typedef T Q __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int W = 16 / (int)sizeof(T);
  const GLB Q* src = gp((const Q*)s.st) + (long)s.n * (TOPT_MG_ST / W);
#pragma unroll
  for (int i = 0; i < TOPT_MG_ST / W; ++i) {
    const Q v = src[i];
#pragma unroll
    for (int w = 0; w < W; ++w) rec[i * W + w] = v[w];
  }
}
// The node's two rows of the level operator applied to x (the level's LDS
// vector); neighbours outside the grid have zero coefficients (read clamped).
// ra(i) gives record entry i; SEQ walks the neighbours one at a time (a
// record read from global memory while the first one stays in registers).
template <typename T, bool SEQ, typename RA>
__device__ __forceinline__ d2v tail_apply(const TailSlot& s, RA ra, const vec2<T>* x) {
  double ax = 0.0, ay = 0.0;
  auto step = [&](int q) {
    const vec2<T> v = x[nbr_node(s.ix, s.iy, q / 3 - 1, q % 3 - 1, s.nx, s.ny)];
    ax = fma((double)ra(2 * q), (double)v.x, ax);
    ax = fma((double)ra(2 * q + 1), (double)v.y, ax);
    ay = fma((double)ra(18 + 2 * q), (double)v.x, ay);
    ay = fma((double)ra(18 + 2 * q + 1), (double)v.y, ay);
  };
  if constexpr (SEQ) {
#pragma unroll 1
    for (int q = 0; q < 9; ++q) step(q);
  } else {
#pragma unroll
    for (int q = 0; q < 9; ++q) step(q);
  }
  return d2v{ax, ay};
}

template <typename T, bool DIRECT>
__global__ __launch_bounds__(TOPT_MG_TAIL_BLOCK) void k_mg_tail(const CgState* cgd, const ProbDev* tab, const int2* map,
                                                                int par) {
  constexpr int NB = TOPT_MG_TAIL_BLOCK;
  constexpr bool REG = sizeof(T) == 4;
  constexpr int G = NB / TOPT_MG_DENSE;   // thread groups of the dense solve
  constexpr int JD = TOPT_MG_DENSE / G;   // its columns per thread
  static_assert(NB % TOPT_MG_DENSE == 0, "tail block: whole groups of coarse rows");
  __shared__ vec2<T> sx[TOPT_MG_TAIL_SUM], sb[TOPT_MG_TAIL_SUM], sr[TOPT_MG_TAIL_SUM];
  static_assert(sizeof(sr) >= sizeof(double) * G * TOPT_MG_DENSE, "dense partial sums fit in sr");
  int lb;
  PD& P = mg_entry(tab, map, blockIdx.x, lb);
  if (mg_stopped<DIRECT>(P, par, cgd)) return;
  const int lt = P.mg_lt, L = P.mg_nlev - 1;
  const int t = (int)threadIdx.x;
  auto nodes = [&](int l) { return (P.mg[l].nx + 1) * (P.mg[l].ny + 1); };
  int tot = 0;   // sweep slots; level L's vectors follow them
  for (int l = lt; l < L; ++l) tot += nodes(l);
  const int nL = nodes(L), N = P.mg_ncd;
  // entry: the first record (FP32)
  const TailSlot s0 = tail_slot(P, lt, L, t < tot ? t : 0);
  T rec0[TOPT_MG_ST];
  if (REG && t < tot) tail_rec_load<T>(s0, rec0);
  const int row = t % TOPT_MG_DENSE, grp = t / TOPT_MG_DENSE;
  // this thread's entries of the coarsest inverse (T: FP32 on the mixed
  // path), fetched with the records, before the down sweeps need the block:
  // A is symmetric, so row `row` is read as column `row` (consecutive
  // threads, consecutive addresses); columns past N read column N-1 and
  // weigh 0 in the solve. With the records in registers (FP32) they are
  // parked in LDS (each thread's own slice) until the solve.
  __shared__ T sinv[REG ? JD * NB : 1];
  T ainv[JD];
  if (row < N) {
    const GLB T* A = gp((const T*)P.mg_inv) + row;
#pragma unroll
    for (int k = 0; k < JD; ++k) {
      const int j = grp + k * G;
      ainv[k] = A[(long)(j < N ? j : N - 1) * N];
    }
  }
  // f(g, slot) for each slot of level l held by this thread
  auto each_slot = [&](int l, auto&& f) {
    for (int g = t; g < tot; g += NB) {
      const TailSlot s = g == t ? s0 : tail_slot(P, lt, L, g);
      if (s.l == l) f(g, s);
    }
  };
  // h(ra) for slot g, ra(i) its record entry i: the registers for the first
  // slot (FP32), else global memory
  auto with_rec = [&](int g, const TailSlot& s, auto&& h) {
    if (REG && g == t) {
      h([&](int i) { return rec0[i]; }, std::integral_constant<bool, false>{});
    } else {
      const GLB T* rp = gp((const T*)s.st) + (long)s.n * TOPT_MG_ST;
      h([=](int i) { return rp[i]; }, std::integral_constant<bool, REG>{});
    }
  };
  auto dinv_of = [&](int g, const TailSlot& s) {
    vec2<T> d;
    if (REG && g == t) {
      d.x = rec0[36];
      d.y = rec0[37];
    } else {
      d = gp((const vec2<T>*)P.mg[s.l].dinv)[s.n];
    }
    return d;
  };
  // b of level lt (formed by the fused sweep of level lt-1) and x_lt
  const GLB vec2<T>* blt = gp((const vec2<T>*)P.mg[lt].b);
  if (lt < L) {
    each_slot(lt, [&](int g, const TailSlot& s) {
      const vec2<T> bb = blt[s.n];
      sb[g] = bb;
      sx[g] = mg_jac<T>(dinv_of(g, s), bb, P.mg_omega);
    });
  } else {
    for (int n = t; n < nL; n += NB) sb[n] = blt[n];
  }
  if constexpr (REG) {
    if (row < N) {
#pragma unroll
      for (int k = 0; k < JD; ++k) sinv[k * NB + t] = ainv[k];
    }
  }
  __syncthreads();
  // down: r_l = b_l - K x_l (x_l = omega D^-1 b_l), then b_{l+1} = P^T r_l
  // and x_{l+1}
  int ol = 0;
  for (int l = lt; l < L; ++l) {
    each_slot(l, [&](int g, const TailSlot& s) {
      with_rec(g, s, [&](auto ra, auto seq) {
        const d2v ax = tail_apply<T, decltype(seq)::value>(s, ra, sx + ol);
        const vec2<T> bn = sb[g];
        vec2<T> ro;
        ro.x = (T)((double)bn.x - ax.x);
        ro.y = (T)((double)bn.y - ax.y);
        sr[g] = ro;
      });
    });
    __syncthreads();
    const vec2<T>* rl = sr + ol;
    if (l + 1 < L) {
      each_slot(l + 1, [&](int g, const TailSlot& s) {
        const vec2<T> bc = mg_restrict_p<T>(P, l, s.n, rl);
        sb[g] = bc;
        sx[g] = mg_jac<T>(dinv_of(g, s), bc, P.mg_omega);
      });
    } else {
      for (int n = t; n < nL; n += NB) sb[tot + n] = mg_restrict_p<T>(P, l, n, rl);
    }
    __syncthreads();
    ol += nodes(l);
  }
  {
    // x_L = A^-1 b_L: thread group grp sums columns j = grp mod G, then the
    // groups' partial sums meet in LDS (in sr, free between the restrictions
    // and the up sweeps)
    double* dpart = (double*)sr;
    const T* bL = (const T*)(sb + tot);
    if (row < N) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < JD; ++k) {
        const int j = grp + k * G;
        const int jc = j < N ? j : N - 1;
        const T a = REG ? sinv[k * NB + t] : ainv[k];
        acc = fma(j < N ? (double)a : 0.0, (double)bL[jc], acc);
      }
      dpart[grp * TOPT_MG_DENSE + row] = acc;
    }
    __syncthreads();
    if (grp == 0 && row < N) {
      double acc = dpart[row];
#pragma unroll
      for (int h = 1; h < G; ++h) acc += dpart[h * TOPT_MG_DENSE + row];
      ((T*)(sx + tot))[row] = (T)acc;
    }
  }
  __syncthreads();
  // up: x' = x + P x_{l+1} in place, then the post-smoothing into r
  for (int l = L - 1; l >= lt; --l) {
    const int oc = ol;   // level l+1's offset
    ol -= nodes(l);
    const vec2<T>* xc = l + 1 == L ? sx + oc : sr + oc;
    const int cny = P.mg[l + 1].ny;
    each_slot(l, [&](int g, const TailSlot& s) {
      const vec2<T> v = sx[g];
      const d2v c = mg_prolong<T>(xc, cny, s.ix, s.iy);
      vec2<T> o;
      o.x = (T)((double)v.x + c.x);
      o.y = (T)((double)v.y + c.y);
      sx[g] = o;
    });
    __syncthreads();
    each_slot(l, [&](int g, const TailSlot& s) {
      with_rec(g, s, [&](auto ra, auto seq) {
        const d2v ax = tail_apply<T, decltype(seq)::value>(s, ra, sx + ol);
        const vec2<T> bn = sb[g], xn = sx[g];
        vec2<T> o;
        o.x = (T)((double)xn.x + P.mg_omega * (double)ra(36) * ((double)bn.x - ax.x));
        o.y = (T)((double)xn.y + P.mg_omega * (double)ra(37) * ((double)bn.y - ax.y));
        sr[g] = o;
      });
    });
    __syncthreads();
  }
  GLB vec2<T>* dst = gp((vec2<T>*)mg_xup(P, lt));
  if (lt < L) {
    each_slot(lt, [&](int g, const TailSlot& s) { dst[s.n] = sr[g]; });
  } else {
    for (int n = t; n < nL; n += NB) dst[n] = sx[n];
  }
}

// Level 0's pre-smoothed iterate at a node from r and 1/diag.
template <typename T>
__device__ __forceinline__ d2v mg_x0(vec2<T> rr, vec2<T> dd, double om) {
  return d2v{om * (double)dd.x * (double)rr.x, om * (double)dd.y * (double)rr.y};
}
// x' = x0 + (masked) P x_1 at the tile's ring node m.
template <typename T>
__device__ __forceinline__ d2v mg_ring_x(PD& P, long m, const GLB vec2<T>* r, const GLB vec2<T>* dv,
                                         const GLB vec2<T>* xc, int cny) {
  const int nny = P.nely + 1;
  const int jx = (int)(m / nny), jy = (int)(m - (long)jx * nny);
  d2v c = mg_prolong<T>(xc, cny, jx, jy);
  const unsigned fb = fixbits(P.fix, m);
  if (fb & 1u) c.x = 0.0;
  if (fb & 2u) c.y = 0.0;
  return mg_x0<T>(r[m], dv[m], P.mg_omega) + c;
}

// Level 0 up sweep: x' = x0 + (masked) P x_1, z = x' + omega D^-1 (r - K x'),
// z = 0 on fixed DOFs; x0 = omega D^-1 r, formed again from r.
template <typename T, bool KST, bool DIRECT>
__global__ __launch_bounds__(TOPT_BLOCK) void k_mg_up0(const CgState* cgd, const ProbDev* tab, const int2* map,
                                                       int par) {
  using Tl = Tile<T>;
  constexpr int NPT = Tl::NPT;
  int lb;
  PD& P = mg_entry(tab, map, xcd_remap(blockIdx.x, gridDim.x), lb);
  if (mg_stopped<DIRECT>(P, par, cgd)) return;
  __shared__ d2v hal[Tl::HN];
  __shared__ double esc[Tl::EN];
  const TileOrg o = tile_org<T>(P, lb);
  const GLB vec2<T>* xc = gp((const vec2<T>*)mg_xup(P, 1));
  const GLB vec2<T>* r = gp((const vec2<T>*)P.r);
  const GLB vec2<T>* dv = gp((const vec2<T>*)P.pre);
  const int cny = P.mg[1].ny;
  TileShare<T> S;
  tile_share<T>(P, o, S);
  d2v hv[NPT];
  vec2<T> orr[NPT], od[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    orr[j] = r[S.ids[j].n];
    od[j] = dv[S.ids[j].n];
    d2v c = mg_prolong<T>(xc, cny, S.ids[j].ix, S.ids[j].iy);
    if (S.fb[j] & 1u) c.x = 0.0;
    if (S.fb[j] & 2u) c.y = 0.0;
    hv[j] = mg_x0<T>(orr[j], od[j], P.mg_omega) + c;
  }
  const d2v rv = mg_ring_x<T>(P, S.rn, r, dv, xc, cny);
  tile_store<T>(S, hv, rv, hal, esc);
  __syncthreads();
  double ax[NPT], ay[NPT];
  tile_apply<T, KST>(P, hal, esc, ax, ay);
  GLB vec2<T>* z = gp((vec2<T>*)P.z);
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    if (!S.ids[j].act) continue;
    vec2<T> zo;
    zo.x = (S.fb[j] & 1u) ? T(0) : (T)(hv[j].x + P.mg_omega * (double)od[j].x * ((double)orr[j].x - ax[j]));
    zo.y = (S.fb[j] & 2u) ? T(0) : (T)(hv[j].y + P.mg_omega * (double)od[j].y * ((double)orr[j].y - ay[j]));
    z[S.ids[j].n] = zo;
  }
}

}  // namespace topt

namespace topt {

// ---------------------------------------------------------------- fused slots
#ifndef TOPT_MG_FUSE_B
#define TOPT_MG_FUSE_B 1   // 0: slot B and the level-0 down sweep as two launches
#endif
// Plans whose problems all use multigrid run slot B with the level-0 down
// sweep and its restriction as one kernel (k_slot_b_mg), and the level-0 up
// sweep with slot C as another (k_slot_c_mg): 11 launches per CG iteration
// instead of 13. The down sweep needs the updated residual around its nodes,
// which neighbouring workgroups of the same launch also update, so r
// alternates between two buffers: iteration parity pa (the parity of its
// slot A) reads r from buffer pa and writes the updated r to buffer pa ^ 1,
// where slot C and the next slot A's residual refresh look for it.

// Level-0 up sweep + slot C: z = x0 + P x_1 + omega D^-1 (r - K x'), z = 0
// on fixed DOFs, partial r.z (x0 = omega D^-1 r, formed again from r).
template <typename T, bool KST>
__global__ __launch_bounds__(TOPT_BLOCK) void k_slot_c_mg(const ProbDev* tab, const int2* pmap, int q, int par,
                                                          int* stat) {
  using Tl = Tile<T>;
  constexpr int NPT = Tl::NPT;
  __shared__ d2v hal[Tl::HN];
  __shared__ double esc[Tl::EN];
  __shared__ double sh[16];
  int lb;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  PD& P = mg_entry(tab, pmap, b, lb);
  const TileOrg o = tile_org<T>(P, lb);
  const GLB vec2<T>* xc = gp((const vec2<T>*)mg_xup(P, 1));
  const GLB vec2<T>* r = gp((const vec2<T>*)(par ? P.r1 : P.r));   // what k_slot_b_mg wrote
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
    orr[j] = r[S.ids[j].n];
    od[j] = dv[S.ids[j].n];
    d2v c = mg_prolong<T>(xc, cny, S.ids[j].ix, S.ids[j].iy);
    if (S.fb[j] & 1u) c.x = 0.0;
    if (S.fb[j] & 2u) c.y = 0.0;
    hv[j] = mg_x0<T>(orr[j], od[j], P.mg_omega) + c;
  }
  const d2v rv = mg_ring_x<T>(P, S.rn, r, dv, xc, cny);
  bool upd;
  const CgState st = slot_begin<T, SK_C>(P, lb, q, pp, sh, &upd, stat, pmap ? pmap[b].x : 0);
  if (st.status != ST_RUN) {
    if (upd) slot_c_stop_update<T>(P, o, st.alpha, par);
    return;
  }
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
    zo.x = (S.fb[j] & 1u) ? T(0) : (T)(hv[j].x + P.mg_omega * (double)od[j].x * ((double)orr[j].x - ax[j]));
    zo.y = (S.fb[j] & 2u) ? T(0) : (T)(hv[j].y + P.mg_omega * (double)od[j].y * ((double)orr[j].y - ay[j]));
    z[S.ids[j].n] = zo;
    v += (double)orr[j].x * (double)zo.x + (double)orr[j].y * (double)zo.y;
  }
  slot_end(P, lb, q, v, 0.0, sh);
}

// r_new = r - alpha Ap of a B_CG step (one rounding per component, the same
// wherever a node's value is formed); a refresh step keeps r.
template <typename T>
__device__ __forceinline__ vec2<T> mg_rnew(vec2<T> r, vec2<T> a, T alpha, bool cg) {
  vec2<T> o;
  o.x = cg ? fma(-alpha, a.x, r.x) : r.x;
  o.y = cg ? fma(-alpha, a.y, r.y) : r.y;
  return o;
}

// Slot B + level-0 down sweep + restriction (fem/fused_kernels.py:110-166
// for the slot part). Workgroup = PCG tile lb (TX = 8 node columns x TY = 64
// node rows from (x0, y0)), so its partial |r|^2 sums have the tile count
// slot C expects (two workgroups per tile, each summing all of slot A's
// partials, measured slower: 13.1 against 11.9 us at 1200x400). par = the
// parity of the p buffer the preceding slot A wrote, as for slot B. Per tile:
//   1. before the state arrives: slot A's partials, r_old (buffer par ^ 1),
//      Ap and 1/diag on the (TX + 3) x (TY + 3) patch from (x0 - 2, y0 - 2)
//      (clamped into the grid), the (TX + 2) x (TY + 2) element scales from
//      (x0 - 2, y0 - 2) (0 outside the grid), the tile's own nodes' r_old
//      and Ap, the residual nodes' fixed bits;
//   2. slot_begin (slot B's state advance: alpha, breakdown), the deferred
//      u += alpha p at the own nodes;
//   3. r_new at the own nodes into buffer par, partial |r_new|^2 (the value
//      slot B adds, in its position);
//   4. x0 = omega D^-1 r_new on the patch (LDS), the fine residuals
//      r_new - K(s) x0 at the (TX + 1) x (TY + 1) nodes from (x0 - 1, y0 - 1)
//      (fixed DOFs and nodes outside the grid: 0), then b_1 = P^T res at the
//      coarse nodes (I, J) with x0 <= 2I < x0 + TX, y0 <= 2J < y0 + TY; the
//      last tile of a row or column also takes the coarse node one further
//      (2I = nelx + 1 for odd nelx), whose fine neighbours past the tile lie
//      outside the grid and add 0.
template <typename T, bool KST>
__global__ __launch_bounds__(TOPT_BLOCK) void k_slot_b_mg(const ProbDev* tab, const int2* pmap, int q, int par) {
  using Tl = Tile<T>;
  constexpr int NPT = Tl::NPT;
  constexpr int TX = Tl::TX, TY = Tl::TY;        // fine nodes of the tile
  constexpr int CX = TX / 2, CY = TY / 2;        // coarse nodes it owns (one more at the grid's end)
  constexpr int RX = TX + 1, RY = TY + 1;        // fine residual nodes
  constexpr int XX = TX + 3, XY = TY + 3;        // x0 patch
  constexpr int EX = TX + 2, EY = TY + 2;        // element scales
  constexpr int NX = (XX * XY + TOPT_BLOCK - 1) / TOPT_BLOCK;
  constexpr int NR = (RX * RY + TOPT_BLOCK - 1) / TOPT_BLOCK;
  constexpr int NE = (EX * EY + TOPT_BLOCK - 1) / TOPT_BLOCK;
  static_assert(TX % 2 == 0 && TY % 2 == 0, "tiles start on coarse nodes");
  static_assert((CX + 1) * (CY + 1) <= TOPT_BLOCK, "one coarse node per thread");
  __shared__ d2v xs[XX * XY];
  __shared__ d2v rs[RX * RY];
  __shared__ double es[EX * EY];
  __shared__ double sh[16];
  int lb;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  PD& P = mg_entry(tab, pmap, b, lb);
  const TileOrg o = tile_org<T>(P, lb);
  const int nelx = P.nelx, nely = P.nely, nny = nely + 1;
  const GLB vec2<T>* rold = gp((const vec2<T>*)(par ? P.r : P.r1));
  GLB vec2<T>* rnew = gp((vec2<T>*)(par ? P.r1 : P.r));
  const GLB vec2<T>* Ap = gp((const vec2<T>*)P.Ap);
  const GLB vec2<T>* dv = gp((const vec2<T>*)P.pre);
  const GLB T* sc = gp((const T*)P.s);
  const int t = (int)threadIdx.x;
  // 1. state-independent loads
  PrevParts pp;
  parts_fetch(P, q, pp);
  vec2<T> pr[NX], pa[NX], pd[NX];
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    const int u0 = t + k * TOPT_BLOCK;
    const int u = u0 < XX * XY ? u0 : XX * XY - 1;
    const int px = u / XY, py = u - (u / XY) * XY;
    int jx = o.x0 - 2 + px, jy = o.y0 - 2 + py;
    jx = jx < 0 ? 0 : (jx > nelx ? nelx : jx);
    jy = jy < 0 ? 0 : (jy > nely ? nely : jy);
    const long m = (long)jx * nny + jy;
    pr[k] = rold[m];
    pa[k] = Ap[m];
    pd[k] = dv[m];
  }
  T pe[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int u0 = t + k * TOPT_BLOCK;
    const int u = u0 < EX * EY ? u0 : EX * EY - 1;
    const int ex = o.x0 - 2 + u / EY, ey = o.y0 - 2 + (u - (u / EY) * EY);
    const bool ok = (ex >= 0) & (ex < nelx) & (ey >= 0) & (ey < nely);
    const int kx = ex < 0 ? 0 : (ex >= nelx ? nelx - 1 : ex), ky = ey < 0 ? 0 : (ey >= nely ? nely - 1 : ey);
    const T v = sc[(long)kx * nely + ky];
    pe[k] = ok ? v : T(0);
  }
  NodeIdx ids[NPT];
  vec2<T> orr[NPT], oa[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    ids[j] = tile_node<T>(P, o, j);
    orr[j] = rold[ids[j].n];
    oa[j] = Ap[ids[j].n];
  }
  unsigned rfb[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int u0 = t + k * TOPT_BLOCK;
    const int u = u0 < RX * RY ? u0 : RX * RY - 1;
    int i = o.x0 - 1 + u / RY, j = o.y0 - 1 + (u - (u / RY) * RY);
    i = i < 0 ? 0 : (i > nelx ? nelx : i);
    j = j < 0 ? 0 : (j > nely ? nely : j);
    rfb[k] = fixbits(P.fix, (long)i * nny + j);
  }
  // 2. the state (slot B's advance)
  bool upd;
  const CgState st = slot_begin<T, SK_B>(P, lb, q, pp, sh, &upd);
  if (upd) {
    // deferred u += alpha p of the last B_CG; p is the buffer slot A wrote (par)
    const GLB vec2<T>* p = gp((const vec2<T>*)(par ? P.p1 : P.p0));
    GLB d2v* u = gp((d2v*)P.u64);
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      if (!ids[j].act) continue;
      const long n = ids[j].n;
      const vec2<T> pj = p[n];
      d2v un = u[n];
      un.x = fma(st.alpha, (double)pj.x, un.x);
      un.y = fma(st.alpha, (double)pj.y, un.y);
      u[n] = un;
    }
  }
  if (st.status != ST_RUN) return;
  const bool cg = st.op == OP_B_CG;
  const T alpha = (T)st.alpha;
  // 3. own nodes
  double v1 = 0.0;
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    if (!ids[j].act) continue;
    const vec2<T> rn = mg_rnew<T>(orr[j], oa[j], alpha, cg);
    rnew[ids[j].n] = rn;
    if (cg) v1 += (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
  }
  // 4. x0 on the patch; r_new at the residual nodes (patch positions 1 .. RX, 1 .. RY)
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    const int u = t + k * TOPT_BLOCK;
    if (u < XX * XY) {
      const vec2<T> rn = mg_rnew<T>(pr[k], pa[k], alpha, cg);
      xs[u] = mg_x0<T>(rn, pd[k], P.mg_omega);
      const int px = u / XY, py = u - (u / XY) * XY;
      if (px >= 1 && px <= RX && py >= 1 && py <= RY) rs[(px - 1) * RY + (py - 1)] = d2v{(double)rn.x, (double)rn.y};
    }
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int u = t + k * TOPT_BLOCK;
    if (u < EX * EY) es[u] = (double)pe[k];
  }
  __syncthreads();
  double kk[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) kk[c] = P.kc[c];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int u = t + k * TOPT_BLOCK;
    if (u < RX * RY) {
      const int qx = u / RY, qy = u - (u / RY) * RY;
      const int i = o.x0 - 1 + qx, j = o.y0 - 1 + qy;
      const bool inr = i >= 0 && i <= nelx && j >= 0 && j <= nely;
      // the node's 3 x 3 neighbourhood at patch positions (qx .. qx + 2, qy .. qy + 2),
      // its four elements at element positions (qx .. qx + 1, qy .. qy + 1)
      d2v nb[3][3];
#pragma unroll
      for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) nb[dx][dy] = xs[(qx + dx) * XY + qy + dy];
      double s4[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) s4[a][c] = es[(qx + a) * EY + qy + c];
      double ax, ay;
      if constexpr (KST) q4_gather_s(kk, s4, nb, ax, ay);
      else q4_gather<double>(P.Ke64, s4, nb, ax, ay);
      const d2v rn = rs[u];
      d2v res = d2v{rn.x - ax, rn.y - ay};
      if (!inr || (rfb[k] & 1u)) res.x = 0.0;
      if (!inr || (rfb[k] & 2u)) res.y = 0.0;
      rs[u] = res;
    }
  }
  __syncthreads();
  const int cnx = P.mg[1].nx, cny = P.mg[1].ny;
  const int cxn = o.x0 + TX > nelx ? CX + 1 : CX, cyn = o.y0 + TY > nely ? CY + 1 : CY;
  if (t < (CX + 1) * (CY + 1)) {
    const int cx = t / (CY + 1), cy = t - (t / (CY + 1)) * (CY + 1);
    const int I = o.x0 / 2 + cx, J = o.y0 / 2 + cy;
    if (cx < cxn && cy < cyn && I <= cnx && J <= cny) {
      double bx = 0.0, by = 0.0;
#pragma unroll
      for (int di = 0; di < 3; ++di)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          const int qx = 2 * cx + di, qy = 2 * cy + dj;
          // residual nodes past the tile (the extra coarse node) lie outside the grid
          if (qx >= RX || qy >= RY) continue;
          const double w = (di == 1 ? 1.0 : 0.5) * (dj == 1 ? 1.0 : 0.5);
          const d2v r = rs[qx * RY + qy];
          bx = fma(w, r.x, bx);
          by = fma(w, r.y, by);
        }
      vec2<T> ob;
      ob.x = (T)bx;
      ob.y = (T)by;
      gp((vec2<T>*)P.mg[1].b)[(long)I * (cny + 1) + J] = ob;
    }
  }
  slot_end(P, lb, q, 0.0, v1, sh);
}

}  // namespace topt
