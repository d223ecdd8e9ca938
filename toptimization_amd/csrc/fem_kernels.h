#pragma once
// FEM kernels of the SIMP hot path: element scales, preconditioner setup,
// node-centric K(rho) matvec, the PCG slot kernels with device-side
// alpha/beta/convergence, and the compliance sensitivity.
//
// Reference counterparts (src/toptimization/):
//   fem/kernels.py:143-176 compute_matvec_2d      -> q4 gather (no atomics)
//   fem/kernels.py:76-101  compute_diagonal_2d    -> k_precond (Jacobi)
//   fem/preconditioner.py:47-108 apply_and_rz_2d  -> k_slot_c (element-block)
//   fem/fused_kernels.py:21-56,110-187            -> k_slot_a / k_slot_b
//   fem/solver.py:27-167 solve                    -> slot sequence driven by engine.hip (host)
//   fem/kernels.py:234-272 compute_sensitivity_2d -> k_sens
#include "topt_dev.h"

namespace topt {

// ---------------------------------------------------------------- helpers

struct NodeIdx {
  long n;      // node id (clamped to a valid node for inactive lanes)
  int ix, iy;
  bool act;
};

__device__ __forceinline__ NodeIdx node_of(PD& P, int lb) {
  NodeIdx id;
  const int nny = P.nely + 1;
  const long nn = (long)(P.nelx + 1) * nny;
  const long n = (long)lb * TOPT_BLOCK + threadIdx.x;
  id.act = n < nn;
  id.n = id.act ? n : nn - 1;
  id.ix = (int)(id.n / nny);
  id.iy = (int)(id.n - (long)id.ix * nny);
  return id;
}

// Element-block Jacobi (reference fem/preconditioner.py:47-108) in gather
// form: z_n = c_n^-1/2 sum_{e ni n} w_e (Ki (r/sqrt(c))_e)[n], fixed -> 0.
// c = number of elements sharing the node (4 interior, 2 edge, 1 corner).
template <typename T>
__device__ __forceinline__ vec2<T> elemblock_node(PD& P, const NodeIdx& id,
                                                  const GLB vec2<T>* r) {
  const T* s = (const T*)P.s;
  T sc[2][2];
  load_scales<T>(s, id.ix, id.iy, P.nelx, P.nely, sc);
  const T areg = (T)P.alpha_reg;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const T c = sc[a][b] > areg ? sc[a][b] : areg;
      sc[a][b] = sc[a][b] / (c * c);   // zero scale (outside) stays zero
    }
  vec2<T> q[3][3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const int jx = id.ix + dx - 1, jy = id.iy + dy - 1;
      const long m = nbr_node(id.ix, id.iy, dx - 1, dy - 1, P.nelx, P.nely);
      const int cx = (jx == 0 || jx == P.nelx) ? 1 : 2;
      const int cy = (jy == 0 || jy == P.nely) ? 1 : 2;
      const T inv = T(1) / sqrt((T)(cx * cy));
      vec2<T> v = r[m];
      v.x *= inv;
      v.y *= inv;
      q[dx][dy] = v;
    }
  T zx, zy;
  q4_gather<T>(KeSel<T>::ki(P), sc, q, zx, zy);
  const int cx = (id.ix == 0 || id.ix == P.nelx) ? 1 : 2;
  const int cy = (id.iy == 0 || id.iy == P.nely) ? 1 : 2;
  const T inv = T(1) / sqrt((T)(cx * cy));
  zx *= inv;
  zy *= inv;
  const unsigned fb = fixbits(P.fix, id.n);
  vec2<T> z;
  z.x = (fb & 1u) ? T(0) : zx;
  z.y = (fb & 2u) ? T(0) : zy;
  return z;
}

// ---------------------------------------------------------------- kernels

// s_e = E_min + rho_e^p (1 - E_min) from the physical density
// (reference fem/kernels.py:92). Writes s in T and, for the mixed path, FP64.
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_scale(const ProbDev* tab, const int2* bmap,
                                                      ProbDev single, int use_filtered) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, bmap, single, b, lb);
  const long ne = (long)P.nelx * P.nely;
  const long e = (long)lb * TOPT_BLOCK + threadIdx.x;
  if (e >= ne) return;
  const GLB double* rho = gp(use_filtered ? P.rhof : P.rho);
  const double sv = P.Emin + pow(rho[e], P.penal) * (1.0 - P.Emin);
  (gp((T*)P.s))[e] = (T)sv;
  if (P.s64 != nullptr && (void*)P.s64 != P.s) gp(P.s64)[e] = sv;
}

// Preconditioner data from the element scales.
// Jacobi: 1/diag, diag = sum s_e Ke_ii, fixed -> 1, clamp 1e-20 (fem/kernels.py:76-101).
// 2x2 block: inverse of the assembled 2x2 diagonal block, fixed DOF -> identity
// row/column (no reference counterpart).
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_precond(const ProbDev* tab, const int2* bmap,
                                                        ProbDev single, int raw) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, bmap, single, b, lb);
  if (P.precond != PC_JACOBI && P.precond != PC_BLOCK2 && P.precond != PC_MG) return;
  const NodeIdx id = node_of(P, lb);
  if (!id.act) return;
  const T* s = (const T*)P.s;
  T sc[2][2];
  load_scales<T>(s, id.ix, id.iy, P.nelx, P.nely, sc);
  const auto K = KeSel<T>::ke(P);
  T dxx = 0, dyy = 0, dxy = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      const int me = (a == 1) ? ((bb == 1) ? 0 : 3) : ((bb == 1) ? 1 : 2);
      dxx += sc[a][bb] * K[(2 * me) * 8 + 2 * me];
      dyy += sc[a][bb] * K[(2 * me + 1) * 8 + 2 * me + 1];
      dxy += sc[a][bb] * K[(2 * me) * 8 + 2 * me + 1];
    }
  const unsigned fb = fixbits(P.fix, id.n);
  vec2<T> out;
  if (P.precond == PC_JACOBI || P.precond == PC_MG) {   // MG: the level-0 smoother
    dxx = (fb & 1u) ? T(1) : (dxx < T(1e-20) ? T(1e-20) : dxx);
    dyy = (fb & 2u) ? T(1) : (dyy < T(1e-20) ? T(1e-20) : dyy);
    out.x = raw ? dxx : T(1) / dxx;
    out.y = raw ? dyy : T(1) / dyy;
    (gp((vec2<T>*)P.pre))[id.n] = out;
  } else {
    if (fb & 1u) { dxx = T(1); dxy = T(0); }
    if (fb & 2u) { dyy = T(1); dxy = T(0); }
    dxx = dxx < T(1e-20) ? T(1e-20) : dxx;
    dyy = dyy < T(1e-20) ? T(1e-20) : dyy;
    T det = dxx * dyy - dxy * dxy;
    if (!(det > T(0))) { dxy = T(0); det = dxx * dyy; }
    const T id_ = T(1) / det;
    out.x = dyy * id_;
    out.y = dxx * id_;
    (gp((vec2<T>*)P.pre))[id.n] = out;
    (gp((T*)P.pre2))[id.n] = -dxy * id_;
  }
}

// Start of a solve: u[fixed] = 0 (or u = 0 cold / zero load),
// p buffers = 0, CG state reset. warm: 0 cold, 1 warm, 2 warm iff the
// problem has finished at least one OC iteration (fem/solver.py:77-83;
// optimizer/simp.py:160 passes warm_start = it > 1).
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_cg_prepare(const ProbDev* tab, const int2* bmap,
                                                           ProbDev single, int warm, int auto_stop) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, bmap, single, b, lb);
  const bool zero_load = P.fnorm2 < 1e-60;
  const int w = (warm == 2) ? (P.oc != nullptr && P.oc->iter > 0) : warm;
  const NodeIdx id = node_of(P, lb);
  if (id.act) {
    GLB d2v* u = gp((d2v*)P.u64);
    d2v v = u[id.n];
    const unsigned fb = fixbits(P.fix, id.n);
    if (!w || zero_load) { v.x = 0.0; v.y = 0.0; }
    if (fb & 1u) v.x = 0.0;
    if (fb & 2u) v.y = 0.0;
    u[id.n] = v;
    vec2<T> zero;
    zero.x = T(0);
    zero.y = T(0);
    if ((void*)P.x != (void*)P.u64) (gp((vec2<T>*)P.x))[id.n] = zero;
    (gp((vec2<T>*)P.p0))[id.n] = zero;
    (gp((vec2<T>*)P.p1))[id.n] = zero;
  }
  if (lb == 0 && threadIdx.x == 0) {
    CgState st{};
    // auto_stop: a design whose OC loop has met its stopping rule runs no solve
    st.status = zero_load ? ST_ZERO : ((auto_stop && P.oc->done) ? ST_DONE : ST_RUN);
    st.op = OP_START;
    st.first = 1;
    st.cflag = CF_CG;
    st.amode = OP_A_CG;
    P.cg[0] = st;
  }
}

// ---------------------------------------------------------------- PCG slots
// The host launches slot kernels A, B[, C], A, B[, C], ... with sequence
// numbers q = 0, 1, 2, ... Kernel q
//   1. reads the state cg[q & 1] written by kernel q-1,
//   2. sums kernel q-1's block partials part[(q-1) & 1] in a fixed order
//      (every block computes the same totals),
//   3. advances the state machine (pcg_advance: alpha, beta, convergence,
//      residual refresh) -- every block reaches the same decision, block 0
//      stores it in cg[(q+1) & 1],
//   4. executes the operation that decision selects and writes its own block
//      partials to part[q & 1].
// No atomics, no grid barrier, one launch per slot. Loads that do not
// depend on the decision (the partials, z, r, Ap, x, preconditioner data,
// element scales, fixed bits) are issued before step 1 completes, so a slot
// pays about two memory latencies: one for those loads and the state, one
// for the p buffer the state selects.
//
// Work decomposition: one block = one tile of TY = 64 node rows (one per
// lane, contiguous in memory) by TX = 4 N node columns (wave w owns
// columns w, w + 4, ...; N = Npt for the PCG kernels, NptMv for the
// op-level matvec). The stencil kernels stage the tile's node halo
// ((TX + 2) x 66 values) and element halo ((TX + 1) x 65 scales) in LDS:
// every value is fetched from L2/HBM once per tile instead of once per
// neighbour, and all of a thread's fetches are in flight together.

template <typename T, int N = Npt<T>::v>
struct Tile {
  static constexpr int NPT = N;
  static constexpr int NW = TOPT_BLOCK / TOPT_WAVE;        // waves per block
  static constexpr int TY = TOPT_WAVE, TX = NW * NPT;      // tile: rows x columns
  static constexpr int HY = TY + 2, HN = (TX + 2) * HY;    // node halo
  static constexpr int EY = TY + 1, EN = (TX + 1) * EY;    // element halo
  static constexpr int HL = (HN + TOPT_BLOCK - 1) / TOPT_BLOCK;
  static constexpr int EL = (EN + TOPT_BLOCK - 1) / TOPT_BLOCK;
};

// Tiles of a grid (the slot kernels' block count per problem).
template <typename T, int N = Npt<T>::v>
__host__ __device__ inline long pcg_tiles(int nelx, int nely) {
  return (long)((nely + TOPT_WAVE) / TOPT_WAVE) * ((nelx + Tile<T, N>::TX) / Tile<T, N>::TX);
}

struct TileOrg {
  int x0, y0;
};

// Tiles are numbered along y first (the memory order of the nodes).
template <typename T, int N = Npt<T>::v>
__device__ __forceinline__ TileOrg tile_org(PD& P, int lb) {
  const int nty = (P.nely + TOPT_WAVE) / TOPT_WAVE;
  const int tx = lb / nty;
  return TileOrg{tx * Tile<T, N>::TX, (lb - tx * nty) * TOPT_WAVE};
}

// Tile node j of this thread; nodes outside the grid are clamped onto it and
// marked inactive (they store nothing and add nothing).
template <typename T, int N = Npt<T>::v>
__device__ __forceinline__ NodeIdx tile_node(PD& P, TileOrg o, int j) {
  NodeIdx id;
  const int ix = o.x0 + (int)(threadIdx.x >> 6) + Tile<T, N>::NW * j;
  const int iy = o.y0 + (int)(threadIdx.x & 63);
  id.act = ix <= P.nelx && iy <= P.nely;
  id.ix = ix <= P.nelx ? ix : P.nelx;
  id.iy = iy <= P.nely ? iy : P.nely;
  id.n = (long)id.ix * (P.nely + 1) + id.iy;
  return id;
}

// Staging map of a tile: thread (wave w, lane) owns the NPT nodes of tile
// column w + 4j, row lane, and loads them as the interior entries of the
// node halo (position (w + 4j + 1, lane + 1)); the halo ring (two columns and
// two rows, Ring::NN entries) is spread over the first threads. Element
// scales likewise: own entries at (w + 4j + 1, lane + 1) and a ring of
// Ring::NE entries (column 0 and row 0). Every load is issued before the
// first wait, with simple index arithmetic.
template <typename T, int N>
struct Ring {
  static constexpr int NN = 2 * Tile<T, N>::HY + 2 * Tile<T, N>::TX;   // node ring entries
  static constexpr int NE = Tile<T, N>::EY + Tile<T, N>::TX;           // element ring entries
};

// Ring node entry t (< Ring::NN) -> halo position (cx, cy).
template <typename T, int N>
__device__ __forceinline__ void ring_node_pos(int t, int& cx, int& cy) {
  constexpr int HY = Tile<T, N>::HY, TX = Tile<T, N>::TX;
  if (t < HY) { cx = 0; cy = t; }
  else if (t < 2 * HY) { cx = TX + 1; cy = t - HY; }
  else if (t < 2 * HY + TX) { cx = t - 2 * HY + 1; cy = 0; }
  else { cx = t - 2 * HY - TX + 1; cy = HY - 1; }
}
// Ring element entry t (< Ring::NE) -> element-tile position (cx, cy).
template <typename T, int N>
__device__ __forceinline__ void ring_elem_pos(int t, int& cx, int& cy) {
  constexpr int EY = Tile<T, N>::EY;
  if (t < EY) { cx = 0; cy = t; }
  else { cx = t - EY + 1; cy = 0; }
}
// Clamped global node at halo position (cx, cy) (as nbr_node).
__device__ __forceinline__ long halo_node_at(PD& P, TileOrg o, int cx, int cy) {
  int jx = o.x0 - 1 + cx, jy = o.y0 - 1 + cy;
  jx = jx < 0 ? 0 : (jx > P.nelx ? P.nelx : jx);
  jy = jy < 0 ? 0 : (jy > P.nely ? P.nely : jy);
  return (long)jx * (P.nely + 1) + jy;
}
// Scale of the element at element-tile position (cx, cy); zero outside the
// grid (as load_scales).
template <typename T>
__device__ __forceinline__ T halo_scale_at(PD& P, const GLB T* s, TileOrg o, int cx, int cy) {
  const int ex = o.x0 - 1 + cx, ey = o.y0 - 1 + cy;
  const bool ok = (ex >= 0) & (ex < P.nelx) & (ey >= 0) & (ey < P.nely);
  const int kx = ex < 0 ? 0 : (ex >= P.nelx ? P.nelx - 1 : ex);
  const int ky = ey < 0 ? 0 : (ey >= P.nely ? P.nely - 1 : ey);
  const T v = s[(long)kx * P.nely + ky];
  return ok ? v : T(0);
}

// State-independent geometry of a thread's share of a tile: its nodes, their
// fixed bits, the ring node, and the element scales (own + ring), fetched.
template <typename T, int N = Npt<T>::v>
struct TileShare {
  NodeIdx ids[N];
  unsigned fb[N];
  T se[N];
  long rn;          // ring node (clamped; valid for every thread)
  int rcx, rcy;     // its halo position
  int ecx, ecy;     // ring element position
  T rse;            // ring element scale
  bool hasr, hase;
};
// INNER: the tile and its halo lie inside the grid (no clamping, every
// node active); see tile_inner.
template <typename T, int N, bool INNER>
__device__ __forceinline__ void tile_share_at(PD& P, TileOrg o, TileShare<T, N>& S) {
  using Tl = Tile<T, N>;
  const GLB T* s = gp((const T*)P.s);
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int nny = P.nely + 1;
#pragma unroll
  for (int j = 0; j < Tl::NPT; ++j) {
    if constexpr (INNER) {
      const int ix = o.x0 + w + Tl::NW * j, iy = o.y0 + lane;
      S.ids[j].ix = ix;
      S.ids[j].iy = iy;
      S.ids[j].act = true;
      S.ids[j].n = (long)ix * nny + iy;
      S.se[j] = s[(long)ix * P.nely + iy];
    } else {
      S.ids[j] = tile_node<T, N>(P, o, j);
      S.se[j] = halo_scale_at<T>(P, s, o, w + Tl::NW * j + 1, lane + 1);
    }
    S.fb[j] = fixbits(P.fix, S.ids[j].n);
  }
  S.hasr = (int)threadIdx.x < Ring<T, N>::NN;
  ring_node_pos<T, N>(S.hasr ? (int)threadIdx.x : 0, S.rcx, S.rcy);
  S.hase = (int)threadIdx.x < Ring<T, N>::NE;
  ring_elem_pos<T, N>(S.hase ? (int)threadIdx.x : 0, S.ecx, S.ecy);
  if constexpr (INNER) {
    S.rn = (long)(o.x0 - 1 + S.rcx) * nny + (o.y0 - 1 + S.rcy);
    S.rse = s[(long)(o.x0 - 1 + S.ecx) * P.nely + (o.y0 - 1 + S.ecy)];
  } else {
    S.rn = halo_node_at(P, o, S.rcx, S.rcy);
    S.rse = halo_scale_at<T>(P, s, o, S.ecx, S.ecy);
  }
}
// True when every halo node and element of the tile lies inside the grid.
template <typename T, int N>
__device__ __forceinline__ bool tile_inner(PD& P, TileOrg o) {
  return o.x0 >= 1 && o.y0 >= 1 && o.x0 + Tile<T, N>::TX <= P.nelx && o.y0 + Tile<T, N>::TY <= P.nely;
}
// N is deduced from S.
template <typename T, int N>
__device__ __forceinline__ void tile_share(PD& P, TileOrg o, TileShare<T, N>& S) {
  if (tile_inner<T, N>(P, o)) tile_share_at<T, N, true>(P, o, S);
  else tile_share_at<T, N, false>(P, o, S);
}

// Write the tile's halo values (own + ring, FP64) and element scales to LDS.
template <typename T, int N>
__device__ __forceinline__ void tile_store(const TileShare<T, N>& S, const d2v (&own)[N], d2v ring,
                                           d2v* hal, double* esc) {
  using Tl = Tile<T, N>;
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
#pragma unroll
  for (int j = 0; j < Tl::NPT; ++j) {
    const int c = w + Tl::NW * j;
    hal[(c + 1) * Tl::HY + lane + 1] = own[j];
    esc[(c + 1) * Tl::EY + lane + 1] = (double)S.se[j];
  }
  if (S.hasr) hal[S.rcx * Tl::HY + S.rcy] = ring;
  if (S.hase) esc[S.ecx * Tl::EY + S.ecy] = (double)S.rse;
}

// K(s) q at the thread's NPT tile nodes from the LDS halo (values and
// scales held in FP64): FP64 Ke, FP64 accumulation. Element-outer order: the
// two Ke rows of our local node in element (a, b) (16 scalars in SGPRs) serve
// all NPT nodes before the next element's rows are loaded.
template <typename T, int N>
__device__ __forceinline__ void tile_stencil(PD& P, const d2v* hal, const double* esc,
                                             double (&ax)[N], double (&ay)[N]) {
  using Tl = Tile<T, N>;
  constexpr int CX[4] = {0, 1, 1, 0};
  constexpr int CY[4] = {0, 0, 1, 1};
  const int w = (int)(threadIdx.x >> 6), row = (int)(threadIdx.x & 63);
#pragma unroll
  for (int j = 0; j < Tl::NPT; ++j) ax[j] = ay[j] = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      // local index of our node inside element (ix - 1 + a, iy - 1 + b)
      const int me = (a == 1) ? ((b == 1) ? 0 : 3) : ((b == 1) ? 1 : 2);
      const auto K0 = P.Ke64 + (2 * me) * 8;
      const auto K1 = K0 + 8;
#pragma unroll
      for (int j = 0; j < Tl::NPT; ++j) {
        const int c = w + Tl::NW * j;
        double tx = 0.0, ty = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const d2v v = hal[(c + a + CX[k]) * Tl::HY + row + b + CY[k]];
          // one FMA per product (a*x + b*y added to t would cost a mul, an fma and an add)
          tx = fma(K0[2 * k], v.x, tx);
          tx = fma(K0[2 * k + 1], v.y, tx);
          ty = fma(K1[2 * k], v.x, ty);
          ty = fma(K1[2 * k + 1], v.y, ty);
        }
        const double sc = esc[(c + a) * Tl::EY + row + b];
        ax[j] = fma(sc, tx, ax[j]);
        ay[j] = fma(sc, ty, ay[j]);
      }
    }
}

// K(s) q at the thread's NPT tile nodes, structured Ke (six values in
// scalar registers), node-outer: each neighbour value is read from LDS once.
template <typename T, int N>
__device__ __forceinline__ void tile_stencil_s(PD& P, const d2v* hal, const double* esc,
                                               double (&ax)[N], double (&ay)[N]) {
  using Tl = Tile<T, N>;
  const int w = (int)(threadIdx.x >> 6), row = (int)(threadIdx.x & 63);
  double k[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) k[t] = P.kc[t];
#pragma unroll
  for (int j = 0; j < Tl::NPT; ++j) {
    const int c = w + Tl::NW * j;
    d2v q[3][3];
    double sc[2][2];
#pragma unroll
    for (int dx = 0; dx < 3; ++dx)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) q[dx][dy] = hal[(c + dx) * Tl::HY + row + dy];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) sc[a][b] = esc[(c + a) * Tl::EY + row + b];
    q4_gather_s(k, sc, q, ax[j], ay[j]);
  }
}

// N is deduced from the result arrays.
template <typename T, bool KST, int N>
__device__ __forceinline__ void tile_apply(PD& P, const d2v* hal, const double* esc, double (&ax)[N],
                                           double (&ay)[N]) {
  if constexpr (KST) tile_stencil_s<T, N>(P, hal, esc, ax, ay);
  else tile_stencil<T, N>(P, hal, esc, ax, ay);
}

// Previous slot's partials, fetched into registers before the state is read.
#define TOPT_PART_PT 8
struct PrevParts {
  d2v v[TOPT_PART_PT];
};
__device__ __forceinline__ void parts_fetch(PD& P, int q, PrevParts& pp) {
  const int nb = P.nb[G_PCG];
  const GLB d2v* part = (const GLB d2v*)(gp((const double*)P.part) + (long)((q - 1) & 1) * nb * 2);
#pragma unroll
  for (int k = 0; k < TOPT_PART_PT; ++k) {
    const int j = (int)threadIdx.x + k * TOPT_BLOCK;
    pp.v[k] = part[j < nb ? j : 0];   // entries j >= nb are masked when summed
  }
}

__device__ __forceinline__ void refresh_done(CgState& s, double rzn) {
  s.beta = s.first ? 0.0 : rzn / (s.rz_prev + 1e-60);
  s.rz = rzn;
  s.first = 0;
  s.rmax2 = s.rr_true;
  s.refines += 1;
  s.op = OP_A_CG;
}

// State machine. In: s.op = what kernel q-1 did, t = its reduced partials.
// Out: s.op = what kernel q does (or status != RUN). Follows fem/solver.py:
// alpha = rz/pAp (:117), breakdown |pAp| < 1e-30 (:114), convergence on the
// recursive |r|^2 < tol^2 |f|^2 (:147), beta = rz/(rz_old + 1e-60) (:151).
// Mixed path: convergence, max_iter and breakdown are confirmed on the FP64
// true residual (A_REF/B_REF), which also replaces the FP32 recursive one
// whenever that has fallen by delta (reliable update).
template <typename T>
__device__ __forceinline__ void pcg_advance(PD& P, CgState& s, const double (&t)[2]) {
  constexpr bool MIXED = sizeof(T) == 4;
  const bool eb = P.precond == PC_ELEMBLOCK || P.precond == PC_MG;   // z comes from slot C
  switch (s.op) {
    case OP_START:
      s.op = OP_A_REF;
      return;
    case OP_A_REF:
      s.rr_true = t[0];
      if (!s.first && t[0] < P.tol2) s.pending = ST_CONV;
      s.op = OP_B_REF;
      return;
    case OP_A_CG:
      s.pAp = t[0];
      if (fabs(t[0]) < 1e-30) {
        if (MIXED) {
          s.pending = ST_BREAK;
          s.op = OP_B_REF;
        } else {
          s.status = ST_BREAK;
        }
      } else {
        s.alpha = s.rz / t[0];
        s.op = OP_B_CG;
      }
      return;
    case OP_B_REF:
      if (s.pending) {
        s.status = s.pending;
        return;
      }
      if (eb) {
        s.cflag = CF_REFRESH;
        s.op = OP_C;
        return;
      }
      refresh_done(s, t[0]);
      return;
    case OP_B_CG: {
      const double rr = t[1];
      s.rr = rr;
      s.it += 1;
      int amode = OP_A_CG;
      if (rr < P.tol2) {
        if (!MIXED) { s.status = ST_CONV; return; }
        amode = OP_A_REF;
      } else if (s.it >= P.max_it) {
        if (!MIXED) { s.status = ST_MAXIT; return; }
        s.pending = ST_MAXIT;
        amode = OP_A_REF;
      } else if (MIXED && rr < P.delta2 * s.rmax2) {
        amode = OP_A_REF;
      }
      if (eb) {
        s.amode = amode;
        s.cflag = CF_CG;
        s.op = OP_C;
        return;
      }
      s.beta = t[0] / (s.rz + 1e-60);
      s.rz_prev = s.rz;
      s.rz = t[0];
      s.op = amode;
      return;
    }
    default:   // OP_C
      if (s.cflag == CF_REFRESH) {
        refresh_done(s, t[0]);
        return;
      }
      s.beta = t[0] / (s.rz + 1e-60);
      s.rz_prev = s.rz;
      s.rz = t[0];
      s.op = s.amode;
      return;
  }
}

// Steps 1-3 of a slot kernel. Returns the advanced state (identical in every
// block); block 0 publishes it for the next kernel. The summation order of
// the partials is fixed: thread t adds j = t, t + 256, ... in order, then the
// block sum in lane and wave order. The sum is formed unconditionally (it is
// unused after OP_START or when the problem has stopped) so that the partial
// loads stay ahead of the state load instead of being sunk behind it.
// The solution update u += alpha p of a B_CG step is deferred to the next
// slot A, which holds p at its nodes anyway (slot B then streams only r, Ap
// and M). The flag upend marks it. Slot A applies it in A_CG mode and when
// the solve has stopped; in A_REF mode neighbours read u, so the refresh
// uses u + alpha p on the fly and B_REF applies the update. In a three-slot
// schedule (element-block Jacobi, multigrid) an FP64 stop at a B_CG step
// (convergence, iteration cap) is decided by slot C's advance; slot C then
// applies the step's deferred update itself (no later slot A will). *upd (out): this kernel must apply u += alpha p at
// its own nodes (alpha from the state).
enum SlotKind : int { SK_A = 0, SK_B = 1, SK_C = 2 };
template <typename T, int K>
__device__ __forceinline__ CgState slot_begin(PD& P, int lb, int q, const PrevParts& pp, double* sh,
                                              bool* upd, int* stat = nullptr, int pi = 0) {
  const CST CgState* cgs = cp((const CgState*)P.cg);
  CgState s;
  {
    const CST CgState& c = cgs[q & 1];
    s.rz = c.rz; s.rz_prev = c.rz_prev; s.alpha = c.alpha; s.beta = c.beta;
    s.pAp = c.pAp; s.rr = c.rr; s.rr_true = c.rr_true; s.rmax2 = c.rmax2;
    s.it = c.it; s.status = c.status; s.op = c.op; s.pending = c.pending;
    s.pad3 = 0; s.first = c.first; s.refines = c.refines; s.amode = c.amode;
    s.cflag = c.cflag; s.upend = c.upend; s.pad1 = s.pad2 = 0;
  }
  const bool was_running = s.status == ST_RUN;
  // keep every load issued so far (partials, state, the caller's
  // state-independent data) ahead of the first wait
  __builtin_amdgcn_sched_barrier(0);
  const int nb = P.nb[G_PCG];
  double t[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < TOPT_PART_PT; ++k) {
    const bool in = (int)threadIdx.x + k * TOPT_BLOCK < nb;
    t[0] += in ? pp.v[k].x : 0.0;
    t[1] += in ? pp.v[k].y : 0.0;
  }
  if (nb > TOPT_PART_PT * TOPT_BLOCK) {   // very large grids: the rest in a loop
    const GLB d2v* part = (const GLB d2v*)(gp((const double*)P.part) + (long)((q - 1) & 1) * nb * 2);
    for (int j = (int)threadIdx.x + TOPT_PART_PT * TOPT_BLOCK; j < nb; j += TOPT_BLOCK) {
      const d2v v = part[j];
      t[0] += v.x;
      t[1] += v.y;
    }
  }
  block_sum<2>(t, sh);
  if (threadIdx.x == 0) {
    sh[0] = t[0];
    sh[1] = t[1];
  }
  __syncthreads();
  t[0] = sh[0];
  t[1] = sh[1];
  if (was_running) pcg_advance<T>(P, s, t);
  const bool pend = was_running && s.upend;
  const bool run = s.status == ST_RUN;
  if (K == SK_A) {
    *upd = pend && !(run && s.op == OP_A_REF);
    s.upend = (pend && !*upd) ? 1 : 0;
  } else if (K == SK_B) {
    *upd = pend && !(run && s.op == OP_B_CG);
    s.upend = (run && s.op == OP_B_CG) ? 1 : 0;
  } else {
    *upd = pend && !run;
    if (*upd) s.upend = 0;
  }
  if (lb == 0 && threadIdx.x == 0) {
    P.cg[(q + 1) & 1] = s;
    // the last slot of a captured chunk also writes the host's status
    // snapshot (mapped pinned memory, read after the chunk's event)
    if (stat != nullptr) {
      stat[2 * pi] = s.status;
      stat[2 * pi + 1] = s.it;
      __threadfence_system();
    }
  }
  return s;
}

// Step 4 tail: this block's partials (two values) for kernel q + 1.
__device__ __forceinline__ void slot_end(PD& P, int lb, int q, double v0, double v1, double* sh) {
  double v[2] = {v0, v1};
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) {
    GLB d2v* part = (GLB d2v*)(gp(P.part) + (long)(q & 1) * P.nb[G_PCG] * 2);
    part[lb] = d2v{v[0], v[1]};
  }
}

// A_REF: r = f - A (u [+ alpha p]) in FP64 (stored in T); partial |r|^2
// (FP64). With a deferred update (upd) the neighbours' u + alpha p is formed
// on the fly. Rare (a few times per solve): direct loads.
template <typename T>
__device__ __forceinline__ double slot_refresh_residual(PD& P, const NodeIdx& id, bool upd, double alpha,
                                                        const GLB vec2<T>* pold, GLB vec2<T>* rdst) {
  const GLB d2v* u = gp((const d2v*)P.u64);
  d2v q[3][3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const long m = nbr_node(id.ix, id.iy, dx - 1, dy - 1, P.nelx, P.nely);
      d2v v = u[m];
      if (upd) {
        const vec2<T> c = pold[m];
        v.x = fma(alpha, (double)c.x, v.x);
        v.y = fma(alpha, (double)c.y, v.y);
      }
      q[dx][dy] = v;
    }
  double sc[2][2];
  load_scales<double>(P.s64, id.ix, id.iy, P.nelx, P.nely, sc);
  double ax, ay;
  q4_gather<double>(P.Ke64, sc, q, ax, ay);
  const unsigned fb = fixbits(P.fix, id.n);
  if (fb & 1u) ax = q[1][1].x;
  if (fb & 2u) ay = q[1][1].y;
  const d2v f = gp((const d2v*)P.f)[id.n];
  const double rx = f.x - ax, ry = f.y - ay;
  if (!id.act) return 0.0;
  vec2<T> r;
  r.x = (T)rx;
  r.y = (T)ry;
  rdst[id.n] = r;
  return rx * rx + ry * ry;
}

// Slot A: A_CG = p = z + beta p (ping-pong buffers, recomputed over the
// halo), Ap = K(s) p, partial p.Ap (fem/fused_kernels.py:21-56 with
// fem/solver.py:151-154 fused in); A_REF = residual refresh, and p carried
// over unchanged into the other buffer. Every slot-A launch of a solve thus
// swaps the p buffers, so the host passes the parity par (p_old = par ? p1 : p0)
// and the p halo is fetched together with z, before the state is known. The
// refreshed residual goes to the residual buffer of the same parity (r1 when
// par is set), which the next slot B reads (one buffer unless the plan fuses
// slot B with the multigrid down sweep).
// FP32 storage, FP64 registers: the stencil of the stored p is summed with
// the FP64 Ke, so the inner operator is K(s32) up to the final rounding of
// Ap (an FP32 Ke and FP32 sums break the rigid-body cancellation and stall
// the FP64 residual refresh).
template <typename T, bool B, bool KST>
__global__ __launch_bounds__(TOPT_BLOCK) void k_slot_a(const ProbDev* tab, const int2* pmap,
                                                       ProbDev single, int q, int par) {
  using Tl = Tile<T>;
  constexpr int NPT = Tl::NPT;
  __shared__ d2v hal[Tl::HN];
  __shared__ double esc[Tl::EN];
  __shared__ double sh[16];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, pmap, single, b, lb);
  const TileOrg o = tile_org<T>(P, lb);
  const GLB vec2<T>* pold = gp((const vec2<T>*)(par ? P.p1 : P.p0));
  GLB vec2<T>* pnew = gp((vec2<T>*)(par ? P.p0 : P.p1));
  const GLB vec2<T>* z = gp((const vec2<T>*)P.z);
  // state-independent loads first
  PrevParts pp;
  parts_fetch(P, q, pp);
  TileShare<T> S;
  tile_share<T>(P, o, S);
  GLB d2v* u = gp((d2v*)P.u64);
  vec2<T> oz[NPT], op[NPT];
  d2v ou[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    oz[j] = z[S.ids[j].n];
    op[j] = pold[S.ids[j].n];
    ou[j] = u[S.ids[j].n];
  }
  const vec2<T> rz = z[S.rn], rp = pold[S.rn];
  bool upd;
  const CgState st = slot_begin<T, SK_A>(P, lb, q, pp, sh, &upd);
  if (upd) {
    // deferred u += alpha p of the last B_CG (p_old is that step's p)
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      if (!S.ids[j].act) continue;
      d2v un = ou[j];
      un.x = fma(st.alpha, (double)op[j].x, un.x);
      un.y = fma(st.alpha, (double)op[j].y, un.y);
      u[S.ids[j].n] = un;
    }
  }
  if (st.status != ST_RUN) return;
  double v = 0.0;
  if (st.op == OP_A_REF) {
    const bool later = st.upend != 0;   // B_REF applies it; form it on the fly
    GLB vec2<T>* rdst = gp((vec2<T>*)(par ? P.r1 : P.r));
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      v += slot_refresh_residual<T>(P, S.ids[j], later, st.alpha, pold, rdst);
      if (S.ids[j].act) pnew[S.ids[j].n] = op[j];
    }
  } else {
    const T beta = (T)st.beta;
    vec2<T> pn[NPT];
    d2v hv[NPT];
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      pn[j].x = fma(beta, op[j].x, oz[j].x);
      pn[j].y = fma(beta, op[j].y, oz[j].y);
      hv[j] = d2v{(double)pn[j].x, (double)pn[j].y};
    }
    vec2<T> rw;
    rw.x = fma(beta, rp.x, rz.x);
    rw.y = fma(beta, rp.y, rz.y);
    tile_store<T>(S, hv, d2v{(double)rw.x, (double)rw.y}, hal, esc);
    __syncthreads();
    double ax[NPT], ay[NPT];
    tile_apply<T, KST>(P, hal, esc, ax, ay);
    GLB vec2<T>* Ap = gp((vec2<T>*)P.Ap);
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      if (S.fb[j] & 1u) ax[j] = (double)pn[j].x;
      if (S.fb[j] & 2u) ay[j] = (double)pn[j].y;
      if (S.ids[j].act) {
        pnew[S.ids[j].n] = pn[j];
        vec2<T> ov;
        ov.x = (T)ax[j];
        ov.y = (T)ay[j];
        Ap[S.ids[j].n] = ov;
        v += (double)pn[j].x * (double)ov.x + (double)pn[j].y * (double)ov.y;
      }
    }
  }
  slot_end(P, lb, q, v, 0.0, sh);
}

// z = M^-1 r at one node from preloaded preconditioner data.
template <typename T>
__device__ __forceinline__ vec2<T> precond_apply(int pc, vec2<T> d, T o, vec2<T> r) {
  vec2<T> z;
  if (pc == PC_JACOBI) {
    z.x = r.x * d.x;
    z.y = r.y * d.y;
  } else if (pc == PC_BLOCK2) {
    z.x = d.x * r.x + o * r.y;
    z.y = o * r.x + d.y * r.y;
  } else {
    z = r;
  }
  return z;
}

// Slot B. B_CG: x += alpha p, r -= alpha Ap, z = M^-1 r, partials r.z, r.r
// (fem/fused_kernels.py:110-166; element-block Jacobi leaves z to slot C as
// fem/fused_kernels.py:169-180 + preconditioner.py do). B_REF: z = M^-1 r,
// partial r.z. The iterate x is the FP64 solution u on both paths: with an
// FP32 iterate the small late updates a p are rounded away against large
// intermediate values and the true residual stops falling (measured on the
// CPU restatement: a 608x152 MBB stalls at 9e-6 with an FP32 iterate and
// converges in the FP64 iteration count with an FP64 one).
template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_slot_b(const ProbDev* tab, const int2* pmap,
                                                       ProbDev single, int q, int par, int* stat) {
  using Tl = Tile<T>;
  __shared__ double sh[16];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, pmap, single, b, lb);
  const TileOrg o = tile_org<T>(P, lb);
  const int pc = P.precond;
  GLB vec2<T>* r = gp((vec2<T>*)P.r);
  GLB vec2<T>* z = gp((vec2<T>*)P.z);
  const GLB vec2<T>* Ap = gp((const vec2<T>*)P.Ap);
  // state-independent loads first
  PrevParts pp;
  parts_fetch(P, q, pp);
  NodeIdx ids[Tl::NPT];
  vec2<T> rv[Tl::NPT], av[Tl::NPT], dv[Tl::NPT];
  T ov[Tl::NPT];
#pragma unroll
  for (int j = 0; j < Tl::NPT; ++j) {
    ids[j] = tile_node<T>(P, o, j);
    const long n = ids[j].n;
    rv[j] = r[n];
    av[j] = Ap[n];
    dv[j] = vec2<T>{T(0), T(0)};
    ov[j] = T(0);
    if (pc == PC_JACOBI || pc == PC_BLOCK2) dv[j] = (gp((const vec2<T>*)P.pre))[n];
    if (pc == PC_BLOCK2) ov[j] = (gp((const T*)P.pre2))[n];
  }
  bool upd;
  const CgState st = slot_begin<T, SK_B>(P, lb, q, pp, sh, &upd, stat, B && pmap ? pmap[b].x : 0);
  if (upd) {
    // deferred u += alpha p of the last B_CG (after a refresh, or at the stop);
    // p is the buffer the preceding slot A wrote (par)
    const GLB vec2<T>* p = gp((const vec2<T>*)(par ? P.p1 : P.p0));
    GLB d2v* u = gp((d2v*)P.u64);
#pragma unroll
    for (int j = 0; j < Tl::NPT; ++j) {
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
  const bool eb = pc == PC_ELEMBLOCK || pc == PC_MG;
  double v0 = 0.0, v1 = 0.0;
  if (st.op == OP_B_REF) {
#pragma unroll
    for (int j = 0; j < Tl::NPT; ++j) {
      if (!ids[j].act) continue;
      const long n = ids[j].n;
      if (!eb) {
        const vec2<T> zz = precond_apply<T>(pc, dv[j], ov[j], rv[j]);
        z[n] = zz;
        v0 += (double)rv[j].x * (double)zz.x + (double)rv[j].y * (double)zz.y;
      }
    }
  } else {
    const T alpha = (T)st.alpha;
#pragma unroll
    for (int j = 0; j < Tl::NPT; ++j) {
      if (!ids[j].act) continue;
      const long n = ids[j].n;
      vec2<T> rn = rv[j];
      rn.x -= alpha * av[j].x;
      rn.y -= alpha * av[j].y;
      r[n] = rn;
      v1 += (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
      if (!eb) {
        const vec2<T> zz = precond_apply<T>(pc, dv[j], ov[j], rn);
        z[n] = zz;
        v0 += (double)rn.x * (double)zz.x + (double)rn.y * (double)zz.y;
      }
    }
  }
  slot_end(P, lb, q, v0, v1, sh);
}

// Slot C (element-block Jacobi and multigrid): z = M^-1 r by gather, partial
// r.z (fem/preconditioner.py:47-108 apply_and_rz_2d); for multigrid the
// V-cycle launched between slots B and C has written z, and slot C forms r.z.
// Slot C at the stop decided by its own advance: the step's deferred
// u += alpha p at the tile's nodes (p: the buffer the preceding slot A wrote,
// par).
template <typename T>
__device__ __forceinline__ void slot_c_stop_update(PD& P, TileOrg o, double alpha, int par) {
  const GLB vec2<T>* p = gp((const vec2<T>*)(par ? P.p1 : P.p0));
  GLB d2v* u = gp((d2v*)P.u64);
#pragma unroll
  for (int j = 0; j < Tile<T>::NPT; ++j) {
    const NodeIdx id = tile_node<T>(P, o, j);
    if (!id.act) continue;
    const vec2<T> pj = p[id.n];
    d2v un = u[id.n];
    un.x = fma(alpha, (double)pj.x, un.x);
    un.y = fma(alpha, (double)pj.y, un.y);
    u[id.n] = un;
  }
}

template <typename T, bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_slot_c(const ProbDev* tab, const int2* pmap,
                                                       ProbDev single, int q, int par, int* stat) {
  using Tl = Tile<T>;
  __shared__ double sh[16];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, pmap, single, b, lb);
  const TileOrg o = tile_org<T>(P, lb);
  PrevParts pp;
  parts_fetch(P, q, pp);
  bool upd;
  const CgState st = slot_begin<T, SK_C>(P, lb, q, pp, sh, &upd, stat, B && pmap ? pmap[b].x : 0);
  if (st.status != ST_RUN) {
    if (upd) slot_c_stop_update<T>(P, o, st.alpha, par);
    return;
  }
  const GLB vec2<T>* r = gp((const vec2<T>*)P.r);
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < Tl::NPT; ++j) {
    const NodeIdx id = tile_node<T>(P, o, j);
    vec2<T> zz;
    if (P.precond == PC_MG) zz = gp((const vec2<T>*)P.z)[id.n];
    else zz = elemblock_node<T>(P, id, r);
    if (!id.act) continue;
    if (P.precond != PC_MG) gp((vec2<T>*)P.z)[id.n] = zz;
    const vec2<T> rv = r[id.n];
    v += (double)rv.x * (double)zz.x + (double)rv.y * (double)zz.y;
  }
  slot_end(P, lb, q, v, 0.0, sh);
}

// Compliance and sensitivities (fem/kernels.py:234-272), FP64:
// dc_e = -p rho^(p-1) (1-E_min) u_e^T Ke u_e, dv_e = 1, c = sum s_e u_e^T Ke u_e.
// rho is the physical density (filtered in density mode).
template <bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_sens(const ProbDev* tab, const int2* bmap,
                                                     ProbDev single, int use_filtered) {
  __shared__ double sh[16];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, bmap, single, b, lb);
  const long ne = (long)P.nelx * P.nely;
  const long e0 = (long)lb * TOPT_BLOCK + threadIdx.x;
  const bool act = e0 < ne;
  const long e = act ? e0 : ne - 1;
  const int ex = (int)(e / P.nely), ey = (int)(e - (long)ex * P.nely);
  const int nny = P.nely + 1;
  const GLB d2v* u = gp((const d2v*)P.u64);
  const long n1 = (long)ex * nny + ey;
  double ue[8];
  const d2v a0 = u[n1], a1 = u[n1 + nny], a2 = u[n1 + nny + 1], a3 = u[n1 + 1];
  ue[0] = a0.x; ue[1] = a0.y; ue[2] = a1.x; ue[3] = a1.y;
  ue[4] = a2.x; ue[5] = a2.y; ue[6] = a3.x; ue[7] = a3.y;
  double uKu = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) t += P.Ke64[i * 8 + j] * ue[j];
    uKu += ue[i] * t;
  }
  const GLB double* rho = gp(use_filtered ? P.rhof : P.rho);
  const double re = rho[e];
  const double sv = P.Emin + pow(re, P.penal) * (1.0 - P.Emin);
  double v[1] = {0.0};
  if (act) {
    gp(P.dc)[e] = -P.penal * pow(re, P.penal - 1.0) * (1.0 - P.Emin) * uKu;
    v[0] = sv * uKu;
  }
  block_sum<1>(v, sh);
  double tot[1];
  if (tree_reduce<1>(v, P.opart, P.ctr[G_ELEM], lb, P.nb[G_ELEM], tot)) {
    if (threadIdx.x == 0) P.oc->c = tot[0];
  }
}

// Op-level K(s) p with the reference BC rule (fixed rows -> identity), and
// optionally p.Ap into out[0] (fem/kernels.py:143-176, fused_kernels.py:21-56).
// The LDS staging of PCG slot A on tiles NptMv<T> nodes per thread wide;
// P.nb[G_PCG] = number of those tiles.
template <typename T, bool KST>
__global__ __launch_bounds__(TOPT_BLOCK) void k_matvec(ProbDev single, const T* __restrict__ pin,
                                                       T* __restrict__ Ap, double* out) {
  constexpr int NPT = NptMv<T>::v;
  using Tl = Tile<T, NPT>;
  PD& P = get_prob<false>(nullptr, nullptr, single, 0);
  __shared__ d2v hal[Tl::HN];
  __shared__ double esc[Tl::EN];
  __shared__ double sh[16];
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const TileOrg o = tile_org<T, NPT>(P, lb);
  const GLB vec2<T>* p = gp((const vec2<T>*)pin);
  TileShare<T, NPT> S;
  tile_share<T, NPT>(P, o, S);
  vec2<T> op[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) op[j] = p[S.ids[j].n];
  const vec2<T> rp = p[S.rn];
  d2v hv[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) hv[j] = d2v{(double)op[j].x, (double)op[j].y};
  tile_store<T, NPT>(S, hv, d2v{(double)rp.x, (double)rp.y}, hal, esc);
  __syncthreads();
  double ax[NPT], ay[NPT];
  tile_apply<T, KST, NPT>(P, hal, esc, ax, ay);
  double v[1] = {0.0};
  GLB vec2<T>* ap = gp((vec2<T>*)Ap);
  vec2<T> ov[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    if (S.fb[j] & 1u) ax[j] = (double)op[j].x;
    if (S.fb[j] & 2u) ay[j] = (double)op[j].y;
    ov[j].x = (T)ax[j];
    ov[j].y = (T)ay[j];
    if (S.ids[j].act) ap[S.ids[j].n] = ov[j];
  }
  if (out == nullptr) return;
#pragma unroll
  for (int j = 0; j < NPT; ++j)
    if (S.ids[j].act) v[0] += (double)op[j].x * (double)ov[j].x + (double)op[j].y * (double)ov[j].y;
  block_sum<1>(v, sh);
  double tot[1];
  if (tree_reduce<1>(v, P.part, P.ctr[G_NODE], lb, P.nb[G_PCG], tot)) {
    if (threadIdx.x == 0) out[0] = tot[0];
  }
}

// Op-level K(s) p for a Ke of the modal form (topt_dev.h), same BC rule and
// optional p.Ap as k_matvec, in the arithmetic of T throughout.
// Each wave owns 64 consecutive node rows y0 - 1 .. y0 + 62 (one per lane)
// and walks the node columns X0 - 1 .. X0 + L. At column x it holds the
// corner values of element (x - 1, y) in registers (the row above through
// a DPP shift from lane + 1, the left column from the previous step), forms
// the element's four node contributions, keeps the two of its own row and
// hands the two of the row above to lane + 1 (DPP shift the other way).
// Node (x - 1, y) is then complete and lanes 1..62 store it. No LDS, no
// barrier, no atomics; each element is formed once per walk (bands overlap
// by two node rows, walks by two node columns). Loads run D columns ahead
// through a register ring (with L + 1 = D every load of the walk is issued
// before the first element). Every lane stays active through the walk (the
// DPP sources); only the stores are predicated.
#ifndef TOPT_MV_D
#define TOPT_MV_D 8
#endif
// FP32 walks read the element scales and write Ap with nontemporal hints
// (neither is used again by the walk): 8192x4096 FP32 122.6 -> 118.1 us;
// the FP64 walk is faster without them (231.7 against 235.8 us).
template <typename T> struct MvNt { static constexpr bool v = sizeof(T) == 4; };
template <typename T>
struct MvWave {
  const GLB vec2<T>* p;
  const GLB T* s;
  const GLB uint32_t* fix;
  GLB vec2<T>* ap;
  int nelx, nely, nny, X0, y, yc, eyc;
  bool erow, slane;
  T c11, c12, c22, c33, hx, hy;
};

// Walk length and prefetch depth (every load of a walk of L + 1 <= TOPT_MV_D
// columns is issued before the first element).
template <int L>
struct MvDepth {
  static constexpr int D = (L + 1) < TOPT_MV_D ? (L + 1) : TOPT_MV_D;
  static_assert((L + 1) % D == 0, "whole prefetch rounds per walk");
};
// Node values of column x and the fixed-bit word of node (x, y) at clamped
// addresses: a node off the grid (clamped onto it, so its value is finite)
// only meets elements off the grid, whose scale is zero. Element scales are
// loaded clamped and zeroed where they are used.
template <typename T>
__device__ __forceinline__ vec2<T> mv_ldp(const MvWave<T>& W, int x) {
  const int xc = x < 0 ? 0 : (x > W.nelx ? W.nelx : x);
  return W.p[(long)xc * W.nny + W.yc];
}
template <typename T>
__device__ __forceinline__ T mv_lds(const MvWave<T>& W, int ex) {
  const int xc = ex < 0 ? 0 : (ex >= W.nelx ? W.nelx - 1 : ex);
  if constexpr (MvNt<T>::v) return __builtin_nontemporal_load(W.s + (long)xc * W.nely + W.eyc);
  else return W.s[(long)xc * W.nely + W.eyc];
}
template <typename T>
__device__ __forceinline__ uint32_t mv_ldf(const MvWave<T>& W, int x) {
  const int xc = x < 0 ? 0 : (x > W.nelx ? W.nelx : x);
  return W.fix[((long)xc * W.nny + W.yc) >> 4];
}

// The walk of one wave from its first D columns (pl: column X0 - 1, pq[d]:
// column X0 + d, sq[d]: element column X0 - 1 + d, loaded by the caller).
// FIX: apply the fixed-row rule from the bitmask, loaded at the step that
// stores the node (the few walks whose bitmask footprint is not all zero;
// the others skip those loads).
template <typename T, int L, bool DOT, bool FIX>
__device__ __forceinline__ double mv_walk(const MvWave<T>& W, vec2<T> (&pq)[MvDepth<L>::D],
                                          T (&sq)[MvDepth<L>::D], vec2<T> pl) {
  // no implicit contraction: every element and node sum rounds the same
  // way at every position of every walk (explicit fma only)
#pragma clang fp contract(off)
  constexpr int D = MvDepth<L>::D;
  const int nelx = W.nelx, nny = W.nny, X0 = W.X0, y = W.y;
  vec2<T> plu;
  plu.x = wave_shift<TOPT_DPP_WAVE_SHL1>(pl.x);
  plu.y = wave_shift<TOPT_DPP_WAVE_SHL1>(pl.y);
  T dlx = plu.x - pl.x, dly = plu.y - pl.y;   // left column: TL - BL
  vec2<T> accL;
  accL.x = accL.y = T(0);
  double dot = 0.0;
  // step k (1 .. L + 1): column x = X0 - 1 + k arrives; element (x - 1, y)
  // is formed; node (x - 1, y) completes (stored from k = 2 on)
  auto step = [&](const int d, const int k, const bool refill) {
    const int x = X0 - 1 + k, ex = x - 1;
    const vec2<T> pc = pq[d];
    const T sc = (W.erow & (ex >= 0) & (ex < nelx)) ? sq[d] : T(0);
    uint32_t fw = 0u;
    if constexpr (FIX) fw = mv_ldf(W, ex);
    if (refill) {
      pq[d] = mv_ldp(W, x + D);
      sq[d] = mv_lds(W, ex + D);
    }
    vec2<T> pu;
    pu.x = wave_shift<TOPT_DPP_WAVE_SHL1>(pc.x);
    pu.y = wave_shift<TOPT_DPP_WAVE_SHL1>(pc.y);
    const T dbx = pc.x - pl.x, dby = pc.y - pl.y;      // BR - BL
    const T dtx = pu.x - plu.x, dty = pu.y - plu.y;    // TR - TL
    const T drx = pu.x - pc.x, dry = pu.y - pc.y;      // TR - BR
    const T A1 = dbx + dtx, A2 = dlx + drx, A3 = dtx - dbx;
    const T B1 = dby + dty, B2 = dly + dry, B3 = dty - dby;
    const T k11 = sc * W.c11, k12 = sc * W.c12, k22 = sc * W.c22, k33 = sc * W.c33;
    const T q1 = fma(k11, A1, k12 * B2), q2 = fma(k12, A1, k22 * B2), q3 = k33 * (A2 + B1);
    const T tx = (sc * W.hx) * A3, ty = (sc * W.hy) * B3;
    const T Ux = q1 + q3, Vx = q1 - q3, Uy = q2 + q3, Vy = q2 - q3;
    accL.x += tx - Ux;
    accL.y += ty - Uy;
    vec2<T> accC;
    accC.x = Vx - tx;
    accC.y = -Vy - ty;
    accC.x += wave_shift<TOPT_DPP_WAVE_SHR1>(Ux + tx);    // TR of lane - 1's element
    accC.y += wave_shift<TOPT_DPP_WAVE_SHR1>(Uy + ty);
    accL.x += wave_shift<TOPT_DPP_WAVE_SHR1>(-Vx - tx);   // TL of lane - 1's element
    accL.y += wave_shift<TOPT_DPP_WAVE_SHR1>(Vy - ty);
    if (k >= 2 && ex <= nelx && W.slane) {   // node (ex, y)
      const long col = (long)ex * nny;
      vec2<T> o = accL;
      if constexpr (FIX) {
        const unsigned fb = (fw >> (2u * (((unsigned)col + (unsigned)y) & 15u))) & 3u;
        if (fb & 1u) o.x = pl.x;
        if (fb & 2u) o.y = pl.y;
      }
      if constexpr (MvNt<T>::v) __builtin_nontemporal_store(o, (W.ap + col) + y);
      else (W.ap + col)[y] = o;
      if constexpr (DOT) dot += (double)pl.x * (double)o.x + (double)pl.y * (double)o.y;
    }
    pl = pc;
    plu = pu;
    dlx = drx;
    dly = dry;
    accL = accC;
  };
  for (int k0 = 1; k0 < L + 1 - D + 1; k0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) step(d, k0 + d, true);
  }
#pragma unroll
  for (int d = 0; d < D; ++d) step(d, L + 2 - D + d, false);
  return dot;
}

#ifndef TOPT_MV_WPE
#define TOPT_MV_WPE 1
#endif
template <typename T, int L, bool DOT>
__global__ __launch_bounds__(TOPT_BLOCK, TOPT_MV_WPE) void k_matvec_modal(ProbDev single, const T* __restrict__ pin,
                                                             T* __restrict__ Ap, double* out) {
  constexpr int D = MvDepth<L>::D;
  PD& P = get_prob<false>(nullptr, nullptr, single, 0);
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  MvWave<T> W;
  W.nelx = P.nelx;
  W.nely = P.nely;
  W.nny = P.nely + 1;
  const int nby = mv_band_groups(W.nely);
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = lb / nby;
  const int yf = ((lb - bx * nby) * (TOPT_BLOCK / TOPT_WAVE) + w) * TOPT_MV_RB - 1;   // lane 0's row
  W.y = yf + lane;
  W.X0 = bx * L;
  W.erow = (W.y >= 0) & (W.y < W.nely);
  W.yc = W.y < 0 ? 0 : (W.y > W.nely ? W.nely : W.y);
  W.eyc = W.y < 0 ? 0 : (W.y >= W.nely ? W.nely - 1 : W.y);
  W.slane = (lane >= 1) & (lane <= TOPT_MV_RB) & (W.y >= 0) & (W.y <= W.nely);
  W.c11 = (T)P.km[0];
  W.c12 = (T)P.km[1];
  W.c22 = (T)P.km[2];
  W.c33 = (T)P.km[3];
  W.hx = (T)P.km[4];
  W.hy = (T)P.km[5];
  W.p = gp((const vec2<T>*)pin);
  W.s = gp((const T*)P.s);
  W.fix = gp(P.fix);
  W.ap = gp((vec2<T>*)Ap);
  // Fixed-bit footprint of the walk: the (at most five) bitmask words of
  // each of its L + 2 columns over the wave's 64 rows, one word per lane;
  // an all-zero footprint (most walks) skips the per-node bit loads. Its
  // load goes first, so the test waits for it alone.
  constexpr bool FOOT = (L + 2) * 5 <= TOPT_WAVE;
  uint32_t fv = 0u;
  bool fok = false;
  if constexpr (FOOT) {
    const int c = lane < (L + 2) * 5 ? lane / 5 : 0, kq = lane - 5 * (lane / 5);
    const int xc = W.X0 - 1 + c < 0 ? 0 : (W.X0 - 1 + c > W.nelx ? W.nelx : W.X0 - 1 + c);
    const int ylo = yf < 0 ? 0 : (yf > W.nely ? W.nely : yf);
    const int yhi = yf + 63 > W.nely ? W.nely : yf + 63;
    const long wlo = ((long)xc * W.nny + ylo) >> 4, whi = ((long)xc * W.nny + yhi) >> 4;
    fok = (lane < (L + 2) * 5) & (wlo + kq <= whi);
    fv = W.fix[fok ? wlo + kq : wlo];   // unconditional load, in bounds
  }
  // the walk's first D columns, in flight before anything waits
  vec2<T> pq[D];
  T sq[D];
  const vec2<T> pl = mv_ldp(W, W.X0 - 1);
#pragma unroll
  for (int d = 0; d < D; ++d) {   // step k = 1 + d: column X0 + d; element and store column X0 - 1 + d
    pq[d] = mv_ldp(W, W.X0 + d);
    sq[d] = mv_lds(W, W.X0 - 1 + d);
  }
  const bool anyfix = FOOT ? __ballot(fok & (fv != 0u)) != 0ull : true;
  const double dot = anyfix ? mv_walk<T, L, DOT, true>(W, pq, sq, pl) : mv_walk<T, L, DOT, false>(W, pq, sq, pl);
  if constexpr (DOT) {
    __shared__ double sh[16];
    double v[1] = {dot};
    block_sum<1>(v, sh);
    double tot[1];
    if (tree_reduce<1>(v, P.part, P.ctr[G_NODE], lb, P.nb[G_PCG], tot)) {
      if (threadIdx.x == 0) out[0] = tot[0];
    }
  }
}

// Op-level element-block Jacobi apply (+ r.z, r.r) for drop-in callers.
template <typename T>
__global__ __launch_bounds__(TOPT_BLOCK) void k_elemblock(ProbDev single, const T* __restrict__ rin,
                                                          T* __restrict__ zout, double* out) {
  PD& P = get_prob<false>(nullptr, nullptr, single, 0);
  __shared__ double sh[16];
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const NodeIdx id = node_of(P, lb);
  const GLB vec2<T>* r = gp((const vec2<T>*)rin);
  const vec2<T> zz = elemblock_node<T>(P, id, r);
  double v[2] = {0.0, 0.0};
  if (id.act) {
    gp((vec2<T>*)zout)[id.n] = zz;
    const vec2<T> rv = r[id.n];
    v[0] = (double)rv.x * (double)zz.x + (double)rv.y * (double)zz.y;
    v[1] = (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
  }
  if (out == nullptr) return;
  block_sum<2>(v, sh);
  double tot[2];
  if (tree_reduce<2>(v, P.part, P.ctr[G_NODE], lb, P.nb[G_NODE], tot)) {
    if (threadIdx.x == 0) {
      out[0] = tot[0];
      out[1] = tot[1];
    }
  }
}

// ---------------------------------------------------------------- BLAS-1
// Op-level vector kernels behind the drop-in API of fem/kernels.py:19-73,
// 330-345 and fem/fused_kernels.py:110-187. One grid-stride kernel, op
// selected by `op`; reductions are FP64 with the last-block pattern.
enum Blas1Op : int {
  B1_FILL = 0,      // x[i] = a                                 (kernels.py:19-23)
  B1_COPY = 1,      // y[i] = x[i]                              (kernels.py:26-30)
  B1_AXPY = 2,      // out[i] = a x[i] + y[i]                    (kernels.py:33-37)
  B1_SCALE = 3,     // out[i] = a x[i]                           (kernels.py:40-44)
  B1_DIV = 4,       // out[i] = x[i] / y[i]                      (kernels.py:65-73)
  B1_DOT = 5,       // red[0] = sum x y                          (kernels.py:47-62)
  B1_MAXDIFF = 6,   // red[0] = max |x - y|                      (kernels.py:339-345)
  B1_SUM = 7,       // red[0] = sum x                            (kernels.py:330-336)
  B1_UPD_JAC = 8,   // u += a p; r -= a Ap; z = r / d; red = (r.z, r.r)  (fused_kernels.py:110-139)
  B1_UPD_NONE = 9,  // u += a p; r -= a Ap; z = r;     red = (r.z, r.r)  (fused_kernels.py:142-166)
  B1_UPD_UR = 10,   // u += a p; r -= a Ap                         (fused_kernels.py:169-180)
  B1_UPD_P = 11,    // p = z + a p                                 (fused_kernels.py:183-187)
  B1_ZERO_FIXED = 12,  // x[i] = 0 where fixed bit i set           (solver.py:162-167)
};

template <typename T>
__global__ __launch_bounds__(TOPT_BLOCK) void k_blas1(int op, long n, double a, T* x, T* y, T* z, T* w,
                                                      T* v, const T* d, const uint32_t* fix, double* part,
                                                      unsigned* counter, double* red) {
  __shared__ double sh[16];
  const long stride = (long)gridDim.x * TOPT_BLOCK;
  const T at = (T)a;
  double acc[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x; i < n; i += stride) {
    switch (op) {
      case B1_FILL: x[i] = at; break;
      case B1_COPY: y[i] = x[i]; break;
      case B1_AXPY: z[i] = at * x[i] + y[i]; break;
      case B1_SCALE: z[i] = at * x[i]; break;
      case B1_DIV: z[i] = x[i] / y[i]; break;
      case B1_DOT: acc[0] += (double)x[i] * (double)y[i]; break;
      case B1_MAXDIFF: {
        const double d = fabs((double)x[i] - (double)y[i]);
        acc[0] = d > acc[0] ? d : acc[0];
        break;
      }
      case B1_SUM: acc[0] += (double)x[i]; break;
      case B1_UPD_JAC:
      case B1_UPD_NONE:
      case B1_UPD_UR: {   // x = u, y = r, z = z, w = p, v = Ap, d = diag
        x[i] += at * w[i];
        const T rv = y[i] - at * v[i];
        y[i] = rv;
        if (op == B1_UPD_UR) break;
        const T zv = op == B1_UPD_JAC ? rv / d[i] : rv;
        z[i] = zv;
        acc[0] += (double)rv * (double)zv;
        acc[1] += (double)rv * (double)rv;
        break;
      }
      case B1_UPD_P: w[i] = z[i] + at * w[i]; break;
      case B1_ZERO_FIXED:
        if ((fix[i >> 5] >> (i & 31)) & 1u) x[i] = T(0);
        break;
    }
  }
  const bool reduce = op == B1_DOT || op == B1_MAXDIFF || op == B1_SUM || op == B1_UPD_JAC || op == B1_UPD_NONE;
  if (!reduce || red == nullptr) return;
  if (op == B1_MAXDIFF) {
    acc[0] = wave_max(acc[0]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = acc[0];
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = sh[0];
      for (int k = 1; k < TOPT_BLOCK / 64; ++k) m = sh[k] > m ? sh[k] : m;
      acc[0] = m;
    }
    __syncthreads();
  } else {
    block_sum<2>(acc, sh);
  }
  double tot[2];
  if (tree_reduce<2>(acc, part, counter, blockIdx.x, gridDim.x, tot, op == B1_MAXDIFF ? 1u : 0u)) {
    if (threadIdx.x == 0) {
      red[0] = tot[0];
      red[1] = tot[1];
    }
  }
}

}  // namespace topt
