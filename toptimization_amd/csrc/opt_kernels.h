#pragma once
// Optimizer kernels of the SIMP hot path: the rmin cone filter as an
// LDS-tiled stencil, and the OC update with an on-device lambda bisection.
//
// Reference counterparts (src/toptimization/):
//   optimizer/filter.py:22-103  precompute_filter   -> host stencil table (taps) +
//                                                      k_filter_tables (padded-table expansion)
//   optimizer/filter.py:106-171 apply_*_filter      -> k_filter (modes)
//   optimizer/oc.py:21-93       oc_update           -> k_oc_reset / k_oc_eval / k_oc_apply
//   fem/kernels.py:330-336      compute_volume      -> folded into k_oc_apply
#include "topt_dev.h"

namespace topt {

// ---------------------------------------------------------------- filter
// out_e = sum_taps w * v(e + tap) over the taps inside the grid, in the
// reference table order (dx outer, dy inner; filter.py:64-72), followed by
// the mode's normalisation. Outside taps read a zero halo, which leaves the
// ordered FP64 sum bit-identical to skipping them.
//   F_HS        v = 1                   out: Hs   = num          (den of filter.py:118-124)
//   F_DENSITY   v = rho                 out: rhof = num / (Hs + 1e-20)   (filter.py:106-124)
//   F_DCAVG     v = dc                  out: dcf  = num / (Hs + 1e-20)   (filter.py:127-149)
//   F_SENS      v = rho dc              out: dcf  = num / (rho Hs + 1e-20) (filter.py:152-171)
//   F_TRANSPOSE v = dc / (Hs + 1e-20)   out: dcf  = num   (exact chain rule H^T, no reference twin)
template <bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_filter(const ProbDev* tab, const int2* bmap,
                                                       ProbDev single, int mode) {
  extern __shared__ double tile[];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, bmap, single, b, lb);
  const int nelx = P.nelx, nely = P.nely, R = P.R;
  const int nty = (nely + TOPT_FILT_TY - 1) / TOPT_FILT_TY;
  const int tx0 = (lb / nty) * TOPT_FILT_TX, ty0 = (lb % nty) * TOPT_FILT_TY;
  const int H = TOPT_FILT_TY + 2 * R, W = TOPT_FILT_TX + 2 * R;
  const GLB double* rho = gp((const double*)P.rho);
  const GLB double* dc = gp((const double*)P.dc);
  const GLB double* Hs = gp((const double*)P.Hs);
  for (int c = threadIdx.x; c < H * W; c += TOPT_BLOCK) {
    const int lx = c / H, ly = c - lx * H;
    const int ex = tx0 - R + lx, ey = ty0 - R + ly;
    double v = 0.0;
    if (ex >= 0 && ex < nelx && ey >= 0 && ey < nely) {
      const long e = (long)ex * nely + ey;
      switch (mode) {
        case F_HS: v = 1.0; break;
        case F_DENSITY: v = rho[e]; break;
        case F_DCAVG: v = dc[e]; break;
        case F_SENS: v = rho[e] * dc[e]; break;
        default: v = dc[e] / (Hs[e] + 1e-20); break;
      }
    }
    tile[c] = v;
  }
  __syncthreads();
  const int ty = threadIdx.x % TOPT_FILT_TY;
  const int txa = threadIdx.x / TOPT_FILT_TY;     // 0..3, second element at txa + 4
  const int ey = ty0 + ty;
  double num0 = 0.0, num1 = 0.0;
  const CST Tap* taps = cp(P.taps);
  const int base0 = (txa + R) * H + ty + R;
  const int base1 = (txa + 4 + R) * H + ty + R;
  for (int k = 0; k < P.ntaps; ++k) {
    const int off = taps[k].dx * H + taps[k].dy;
    const double w = taps[k].w;
    num0 += w * tile[base0 + off];
    num1 += w * tile[base1 + off];
  }
  if (ey >= nely) return;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int ex = tx0 + txa + 4 * h;
    if (ex >= nelx) continue;
    const long e = (long)ex * nely + ey;
    const double num = h ? num1 : num0;
    switch (mode) {
      case F_HS: gp(P.Hs)[e] = num; break;
      case F_DENSITY: gp(P.rhof)[e] = num / (Hs[e] + 1e-20); break;
      case F_DCAVG: gp(P.dcf)[e] = num / (Hs[e] + 1e-20); break;
      case F_SENS: gp(P.dcf)[e] = num / (rho[e] * Hs[e] + 1e-20); break;
      default: gp(P.dcf)[e] = num; break;
    }
  }
}

// Expand the stencil into the reference's padded neighbour tables
// (filter.py:52-74): neighbours (-1 pad), float32 weights (0 pad), counts.
// One thread per element; the row order is the reference's scan order.
__global__ __launch_bounds__(TOPT_BLOCK) void k_filter_tables(int nelx, int nely, const Tap* taps,
                                                              int ntaps, int width, int* nb,
                                                              float* wt, int* cnt) {
  const long e = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x;
  if (e >= (long)nelx * nely) return;
  const int ex = (int)(e / nely), ey = (int)(e - (long)ex * nely);
  int k = 0;
  for (int t = 0; t < ntaps; ++t) {
    const int jx = ex + taps[t].dx, jy = ey + taps[t].dy;
    if (jx < 0 || jx >= nelx || jy < 0 || jy >= nely) continue;
    nb[e * width + k] = jx * nely + jy;
    wt[e * width + k] = (float)taps[t].w;
    ++k;
  }
  for (int j = k; j < width; ++j) {
    nb[e * width + j] = -1;
    wt[e * width + j] = 0.0f;
  }
  cnt[e] = k;
}

// ---------------------------------------------------------------- OC
// Bisection on lambda (optimizer/oc.py:52-93): bracket [0, 1e9], at most 50
// steps, stop when (hi-lo)/(hi+lo+1e-30) < 1e-4, commit rho(last lambda).
// Each k_oc_eval pass evaluates the volumes of 2^D - 1 bisection midpoints
// in one sweep over the elements, then the last block replays bisection steps
// from those volumes. The first pass of an OC step (CHAIN) takes the chain of
// midpoints 1e9 / 2^(k+1), k < 2^D - 1: from lo = 0 every midpoint is an exact
// halving, and the bisection keeps moving hi down until the first volume
// above volfrac, so one sweep covers that whole prefix (about 25 steps for
// typical lambdas) and the replay stops at the first step that moves lo. The
// later passes take a whole depth-D subtree of the current bracket. The
// midpoint sequence is exactly the reference's (mid = 0.5*(lo+hi) in FP64);
// volumes are FP64 sums.

template <bool B>
__global__ void k_oc_reset(const ProbDev* tab, const int* idx, int nprob, ProbDev single) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (B ? nprob : 1)) return;
  const int i = idx ? idx[t] : t;
  PD& P = B ? *(PD*)(tab + i) : *(PD*)&single;
  OcState* o = P.oc;
  o->lo = 0.0;
  o->hi = 1e9;
  o->last_mid = 0.5e9;
  o->steps = 0;
  o->stopped = 0;
}

// Lambda of subtree node j (heap order: children 2j+1 -> "vol <= volfrac"
// (hi = mid), 2j+2 -> "vol > volfrac" (lo = mid)) starting from [lo, hi].
__device__ __forceinline__ double subtree_lambda(int j, double lo, double hi) {
  int path[TOPT_OC_MAXD];
  int d = 0;
  while (j > 0) {
    path[d++] = j;
    j = (j - 1) >> 1;
  }
  double mid = 0.5 * (lo + hi);
  for (int k = d - 1; k >= 0; --k) {
    const int c = path[k];
    if (c & 1) hi = mid;     // left child
    else lo = mid;           // right child
    mid = 0.5 * (lo + hi);
  }
  return mid;
}

// Lambda of chain entry j from [0, hi]: j + 1 halvings (exact).
__device__ __forceinline__ double chain_lambda(int j, double hi) {
  double mid = 0.5 * hi;
  for (int k = 0; k < j; ++k) mid = 0.5 * mid;
  return mid;
}

template <int D, bool B, bool CHAIN = false>
__global__ __launch_bounds__(TOPT_BLOCK) void k_oc_eval(const ProbDev* tab, const int2* bmap,
                                                        ProbDev single) {
  constexpr int NL = (1 << D) - 1;
  __shared__ double qs[NL];
  __shared__ double red[4 * NL];
  __shared__ double vtot[NL];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, bmap, single, b, lb);
  OcState* o = P.oc;
  if (o->stopped) return;
  const double lo = o->lo, hi = o->hi;
  // the chain is only valid from the initial bracket (lo = 0)
  const bool chain = CHAIN && lo == 0.0;
  if (threadIdx.x < NL)
    qs[threadIdx.x] =
        1.0 / sqrt((chain ? chain_lambda(threadIdx.x, hi) : subtree_lambda(threadIdx.x, lo, hi)) + 1e-20);
  __syncthreads();
  const long ne = (long)P.nelx * P.nely;
  const double mv = P.move;
  double acc[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) acc[j] = 0.0;
  const long base = (long)lb * TOPT_BLOCK * TOPT_OC_PER_THREAD + threadIdx.x;
#pragma unroll 2
  for (int k = 0; k < TOPT_OC_PER_THREAD; ++k) {
    const long e = base + (long)k * TOPT_BLOCK;
    if (e >= ne) break;
    const double r = gp(P.rho)[e];
    const double g = -gp(P.dcf)[e];
    const double a = r * sqrt(g > 0.0 ? g : 0.0);
    const double rl = (r - mv) > 0.001 ? (r - mv) : 0.001;
    const double rh = (r + mv) < 1.0 ? (r + mv) : 1.0;
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      double c = a * qs[j];
      c = c < rh ? c : rh;
      acc[j] += (c > rl ? c : rl);
    }
  }
  // block reduction of NL sums (fixed order), then the grid tree
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const double t = wave_sum(acc[j]);
    if (lane == 0) red[wv * NL + j] = t;
  }
  __syncthreads();
  double vsum[NL];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NL; ++j) vsum[j] = red[j] + red[NL + j] + red[2 * NL + j] + red[3 * NL + j];
  }
  double tot[NL];
  if (!tree_reduce<NL>(vsum, P.opart, P.ctr[G_OC], lb, P.nb[G_OC], tot)) return;
  // the replay indexes the volumes with the step's node: from LDS, not a
  // dynamically indexed register array (scratch)
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NL; ++j) vtot[j] = tot[j];
  }
  if (threadIdx.x == 0) {
    double l = lo, h = hi;
    int node = 0, steps = o->steps;
    const double nd = (double)ne;
    for (int d = 0; d < (chain ? NL : D); ++d) {
      const double mid = 0.5 * (l + h);
      const double vol = vtot[node] / nd;
      bool right = false;
      if (vol > P.volfrac) {
        l = mid;
        node = 2 * node + 2;
        right = true;
      } else {
        h = mid;
        node = chain ? d + 1 : 2 * node + 1;
      }
      ++steps;
      o->last_mid = mid;
      if ((h - l) / (h + l + 1e-30) < 1e-4 || steps >= 50) {
        o->stopped = 1;
        break;
      }
      if (chain && right) break;   // off the chain: the subtree passes go on
    }
    o->lo = l;
    o->hi = h;
    o->steps = steps;
  }
}

// Commit rho <- clip(rho sqrt(max(-dcf/(lam dv + 1e-20), 0)), ...) at the last
// evaluated lambda (optimizer/oc.py:40-47, 91-92), max |drho| and the mean
// density (fem/kernels.py:330-336) by the last block.
template <bool B>
__global__ __launch_bounds__(TOPT_BLOCK) void k_oc_apply(const ProbDev* tab, const int2* bmap,
                                                         ProbDev single) {
  __shared__ double red[8];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lb;
  PD& P = grid_prob<B>(tab, bmap, single, b, lb);
  OcState* o = P.oc;
  const long ne = (long)P.nelx * P.nely;
  const double lam = o->last_mid, mv = P.move;
  double vsum = 0.0, vmax = 0.0;
  const long base = (long)lb * TOPT_BLOCK * TOPT_OC_PER_THREAD + threadIdx.x;
  for (int k = 0; k < TOPT_OC_PER_THREAD; ++k) {
    const long e = base + (long)k * TOPT_BLOCK;
    if (e >= ne) break;
    const double r = gp(P.rho)[e];
    const double dv = P.dv ? gp(P.dv)[e] : 1.0;
    double be = -gp(P.dcf)[e] / (lam * dv + 1e-20);
    be = be > 0.0 ? be : 0.0;
    const double cand = r * sqrt(be);
    const double rl = (r - mv) > 0.001 ? (r - mv) : 0.001;
    const double rh = (r + mv) < 1.0 ? (r + mv) : 1.0;
    const double c = cand < rh ? cand : rh;
    const double rn = c > rl ? c : rl;
    gp(P.rho)[e] = rn;
    vsum += rn;
    const double d = fabs(rn - r);
    vmax = d > vmax ? d : vmax;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  vsum = wave_sum(vsum);
  vmax = wave_max(vmax);
  if (lane == 0) {
    red[wv] = vsum;
    red[4 + wv] = vmax;
  }
  __syncthreads();
  double v2[2] = {0.0, 0.0};
  if (threadIdx.x == 0) {
    v2[0] = red[0] + red[1] + red[2] + red[3];
    double m = red[4];
    for (int w = 5; w < 8; ++w) m = red[w] > m ? red[w] : m;
    v2[1] = m;
  }
  double tot[2];
  if (!tree_reduce<2>(v2, P.opart, P.ctr[G_OC], lb, P.nb[G_OC], tot, 2u)) return;
  if (threadIdx.x == 0) {
    const double s = tot[0], m = tot[1];
    o->vol = s / (double)ne;
    o->change = m;
    o->iter += 1;
    // stopping rule of optimizer/simp.py:217 (used by the asynchronous driver)
    o->done = ((o->iter > 5 && m < P.oc_tol) || o->iter >= P.oc_max_iter) ? 1 : 0;
  }
}

// Mean of an element field (fem/kernels.py:330-336) into out[0].
__global__ __launch_bounds__(TOPT_BLOCK) void k_volume(ProbDev single, const double* __restrict__ v,
                                                       double* out) {
  PD& P = get_prob<false>(nullptr, nullptr, single, 0);
  __shared__ double sh[16];
  const int lb = blockIdx.x;
  const long ne = (long)P.nelx * P.nely;
  double acc[1] = {0.0};
  const long base = (long)lb * TOPT_BLOCK * TOPT_OC_PER_THREAD + threadIdx.x;
  for (int k = 0; k < TOPT_OC_PER_THREAD; ++k) {
    const long e = base + (long)k * TOPT_BLOCK;
    if (e < ne) acc[0] += v[e];
  }
  block_sum<1>(acc, sh);
  double tot[1];
  if (tree_reduce<1>(acc, P.opart, P.ctr[G_OC], lb, P.nb[G_OC], tot)) {
    if (threadIdx.x == 0) out[0] = tot[0] / (double)ne;
  }
}

}  // namespace topt
