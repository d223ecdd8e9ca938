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

// An OC candidate clipped to [rl, rh] (upper bound first). Here c >= +0 and
// 0 < rl <= rh, so v_min/v_max give the bits of the compare-and-select form
// at half its instructions.
__device__ __forceinline__ double oc_clip(double c, double rl, double rh) {
  return __builtin_fmax(__builtin_fmin(c, rh), rl);
}

// The bisection steps of one pass from its volumes vtot (optimizer/oc.py:
// 71-88): up to NL steps along the chain (from lo = 0: left steps halve hi,
// the first right step ends the pass) or D steps down the subtree of
// [lo, hi]. Thread j first forms node j's bracket, midpoint, volume
// division and the stopping test of both outcomes, with the operations of a
// serial walk on the same operands (the same bits); thread 0 then walks the
// taken path. Every thread calls it; l, h, last (the last midpoint), steps
// and stop are thread 0's results. sh: 3 NL doubles, sf: NL ints (LDS).
template <int D>
__device__ __forceinline__ void oc_replay(bool chain, double lo, double hi, double volfrac, double nd,
                                          const double* vtot, double* sh, int* sf, double& l, double& h,
                                          double& last, int& steps, int& stop) {
  constexpr int NL = (1 << D) - 1;
  const int t = (int)threadIdx.x;
  if (t < NL) {
    double a = lo, b = hi;   // node t's bracket before its step
    if (chain) {
      for (int k = 0; k < t; ++k) b = 0.5 * (a + b);
    } else {
      int path[TOPT_OC_MAXD];
      int d = 0;
      for (int j = t; j > 0; j = (j - 1) >> 1) path[d++] = j;
      for (int k = d - 1; k >= 0; --k) {
        const double mid = 0.5 * (a + b);
        if (path[k] & 1) b = mid;   // left child: vol <= volfrac
        else a = mid;               // right child: vol > volfrac
      }
    }
    const double mid = 0.5 * (a + b);
    const bool right = vtot[t] / nd > volfrac;
    const double nl = right ? mid : a, nh = right ? b : mid;
    const bool sw = (nh - nl) / (nh + nl + 1e-30) < 1e-4;
    sh[t] = mid;
    sh[NL + t] = nl;
    sh[2 * NL + t] = nh;
    sf[t] = (right ? 1 : 0) | (sw ? 2 : 0);
  }
  __syncthreads();
  if (t == 0) {
    int node = 0;
    stop = 0;
    for (int d = 0; d < (chain ? NL : D); ++d) {
      const int f = sf[node];
      const bool right = (f & 1) != 0;
      last = sh[node];
      l = sh[NL + node];
      h = sh[2 * NL + node];
      ++steps;
      if ((f & 2) || steps >= 50) {
        stop = 1;
        break;
      }
      if (chain && right) break;   // off the chain: subtree passes go on
      node = right ? 2 * node + 2 : (chain ? d + 1 : 2 * node + 1);
    }
  }
}

template <int D, bool B, bool CHAIN = false>
__global__ __launch_bounds__(TOPT_BLOCK) void k_oc_eval(const ProbDev* tab, const int2* bmap,
                                                        ProbDev single) {
  constexpr int NL = (1 << D) - 1;
  __shared__ double qs[NL];
  __shared__ double red[4 * NL];
  __shared__ double vtot[NL];
  __shared__ double rsh[3 * NL];
  __shared__ int rsf[NL];
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
    for (int j = 0; j < NL; ++j) acc[j] += oc_clip(a * qs[j], rl, rh);
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
  __syncthreads();
  double l = lo, h = hi, last = 0.0;
  int steps = o->steps, stop = 0;
  oc_replay<D>(chain, lo, hi, P.volfrac, (double)ne, vtot, rsh, rsf, l, h, last, steps, stop);
  if (threadIdx.x == 0) {
    o->last_mid = last;
    if (stop) o->stopped = 1;
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

// ---------------------------------------------------------------- one-launch OC
// The OC update of a plan (optimizer/oc.py:21-93) in one launch instead of
// reset, chain pass, ten subtree passes and commit. A grid of G blocks, all
// resident at once (G is at most the device's resident capacity for this
// kernel, which the host checks), walks the logical OC blocks of the launch
// (block b holds v = b, b + G, ..., at most TOPT_OCC_MAXK of them). Each pass
// evaluates the lambda set of k_oc_eval (the chain of halvings while the
// bracket starts at 0, else a depth-D subtree of the bracket) over every held
// logical block of a running problem and hands the NL block sums to the
// problem's reduction tree (tree_reduce, as k_oc_eval does). The block that
// completes a problem's tree replays the bisection steps as k_oc_eval does and
// publishes the problem's new bracket in its OcState with a version number;
// the problem's other blocks wait for that version and read the bracket.
// Volumes, lambdas and the step sequence are therefore bit-identical to the
// multi-launch path, and so is the commit, which follows k_oc_apply.
//
// Blocks only wait on problems they hold, after publishing their sums for
// every held problem, and tree_reduce never waits, so every wait ends once
// all blocks are resident. Each wait is bounded by TOPT_OCC_SPIN_TICKS of
// the 100 MHz real-time counter (or the plan's own limit): a block that
// times out raises the plan's abort word, which ends every other wait; rho
// stays unchanged, and while the word is set (it stays set) the plan's
// statistics gather reports oc_steps = -1, which the host turns into an
// error. Versions grow by one per pass and are compared modulo 2^32 from
// the value each block reads on entry.
#define TOPT_OCC_MAXG 256
#define TOPT_OCC_MAXK 16
#define TOPT_OCC_PASSES 12                 // >= 11: one chain step and ten 5-step passes reach 50
#define TOPT_OCC_SPIN_TICKS 100000000ull   // 1 s at 100 MHz: far beyond any wait of a resident grid
#define TOPT_OCC_WORDS 64                  // plan-level words: [0] abort, [2..3] spin limit (0: default)

template <typename X>
__device__ __forceinline__ X occ_get(const X* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename X>
__device__ __forceinline__ void occ_put(X* p, X v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0: wait until the version at ver reaches want. False after a timeout
// of lim ticks (which raises the abort word) or when the abort word is set.
__device__ bool occ_wait(const unsigned* ver, unsigned want, unsigned* abort_w, unsigned long long lim) {
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  while ((int)(occ_get(ver) - want) < 0) {
    if (occ_get(abort_w) != 0u) return false;
    if (__builtin_amdgcn_s_memrealtime() - t0 > lim) {
      occ_put(abort_w, 1u);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  // the bracket loads that follow must not move above the wait
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return true;
}

__device__ __forceinline__ PD& occ_prob(const ProbDev* tab, const int2* bmap, int v, int& lb) {
  if (bmap == nullptr) {
    lb = v;
    return *(PD*)tab;
  }
  const int2 e = bmap[v];
  lb = e.y;
  return *(PD*)(tab + e.x);
}

// Volumes of the pass's NL lambdas over one logical block (the sums of
// k_oc_eval, element order and all), handed to the problem's reduction tree;
// the block that completes the tree replays the bisection and publishes the
// bracket. The per-element factors come from registers (logical block 0,
// loaded once) or from memory.
template <int D, bool CACHED>
__device__ __forceinline__ void occ_pass(PD& P, int lb, int k, unsigned v0, int pass, double* qs, double* red,
                                         double* vtot, double* rsh, int* rsf, double* s_lo, double* s_hi,
                                         double* s_mid, int* s_steps, int* s_stop, double (&ca)[TOPT_OC_PER_THREAD],
                                         double (&crl)[TOPT_OC_PER_THREAD], double (&crh)[TOPT_OC_PER_THREAD]) {
  constexpr int NL = (1 << D) - 1;
  const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
  const double lo = s_lo[k], hi = s_hi[k];
  // the chain (exact halvings of hi) holds while the bracket starts at 0,
  // which k_oc_eval tests the same way
  const bool chain = lo == 0.0;
  if (t < NL) qs[t] = 1.0 / sqrt((chain ? chain_lambda(t, hi) : subtree_lambda(t, lo, hi)) + 1e-20);
  __syncthreads();
  double acc[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) acc[j] = 0.0;
  const long ne = (long)P.nelx * P.nely;
  const long base = (long)lb * TOPT_BLOCK * TOPT_OC_PER_THREAD + t;
  if constexpr (CACHED) {
    // elements past the grid carry a = rl = rh = 0 and add exact zeros
    // after the element sums, which leaves them unchanged
#pragma unroll
    for (int kk = 0; kk < TOPT_OC_PER_THREAD; ++kk)
#pragma unroll
      for (int j = 0; j < NL; ++j) acc[j] += oc_clip(ca[kk] * qs[j], crl[kk], crh[kk]);
  } else {
    const double mv = P.move;
#pragma unroll 2
    for (int kk = 0; kk < TOPT_OC_PER_THREAD; ++kk) {
      const long e = base + (long)kk * TOPT_BLOCK;
      if (e >= ne) break;
      const double r = gp(P.rho)[e];
      const double g = -gp(P.dcf)[e];
      const double a = r * sqrt(g > 0.0 ? g : 0.0);
      const double rl = (r - mv) > 0.001 ? (r - mv) : 0.001;
      const double rh = (r + mv) < 1.0 ? (r + mv) : 1.0;
#pragma unroll
      for (int j = 0; j < NL; ++j) acc[j] += oc_clip(a * qs[j], rl, rh);
    }
  }
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const double tw = wave_sum(acc[j]);
    if (lane == 0) red[wv * NL + j] = tw;
  }
  __syncthreads();
  double vsum[NL];
  if (t == 0) {
#pragma unroll
    for (int j = 0; j < NL; ++j) vsum[j] = red[j] + red[NL + j] + red[2 * NL + j] + red[3 * NL + j];
  }
  double tot[NL];
  if (tree_reduce<NL>(vsum, P.opart, P.ctr[G_OC], lb, P.nb[G_OC], tot)) {
    // this block completed the problem's tree: replay and publish
    if (t == 0) {
#pragma unroll
      for (int j = 0; j < NL; ++j) vtot[j] = tot[j];
    }
    __syncthreads();
    double l = lo, h = hi, last = 0.0;
    int steps = s_steps[k], stop = 0;
    oc_replay<D>(chain, lo, hi, P.volfrac, (double)ne, vtot, rsh, rsf, l, h, last, steps, stop);
    if (t == 0) {
      OcState* o = P.oc;
      occ_put(&o->lo, l);
      occ_put(&o->hi, h);
      occ_put(&o->last_mid, last);
      occ_put(&o->steps, steps);
      occ_put(&o->stopped, stop);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      occ_put(&o->ver, v0 + (unsigned)pass + 1u);
    }
  }
  __syncthreads();   // qs and red are reused by the next logical block
}

template <int D>
__global__ __launch_bounds__(TOPT_BLOCK) void k_oc_coop(const ProbDev* tab, const int2* bmap, int nvb,
                                                        unsigned* bar) {
  constexpr int NL = (1 << D) - 1;
  __shared__ double qs[NL];
  __shared__ double red[4 * NL];
  __shared__ double vtot[NL];
  __shared__ double rsh[3 * NL];
  __shared__ int rsf[NL];
  __shared__ double s_lo[TOPT_OCC_MAXK], s_hi[TOPT_OCC_MAXK], s_mid[TOPT_OCC_MAXK];
  __shared__ int s_steps[TOPT_OCC_MAXK], s_stop[TOPT_OCC_MAXK];
  __shared__ unsigned s_v0[TOPT_OCC_MAXK];
  __shared__ int s_ok;
  __shared__ double ared[8];
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
  const int K = (nvb - b + G - 1) / G;   // logical blocks held (G <= nvb, so K >= 1)
  // the plan may lower the spin limit (tests of the abort path)
  const unsigned long long lset = *(const unsigned long long*)(bar + 2);
  const unsigned long long lim = lset != 0ull ? lset : TOPT_OCC_SPIN_TICKS;
  if (t < K) {
    int lb;
    PD& P = occ_prob(tab, bmap, b + t * G, lb);
    s_v0[t] = occ_get(&P.oc->ver);   // read before any sum of this launch is published
    s_lo[t] = 0.0;
    s_hi[t] = 1e9;
    s_mid[t] = 0.5e9;
    s_steps[t] = 0;
    s_stop[t] = 0;
  }
  if (t == 0) s_ok = 1;
  // per-element factors of logical block 0, loaded once
  double ca[TOPT_OC_PER_THREAD], crl[TOPT_OC_PER_THREAD], crh[TOPT_OC_PER_THREAD];
  int lb0;
  PD& P0 = occ_prob(tab, bmap, b, lb0);
  {
    const long ne = (long)P0.nelx * P0.nely;
    const double mv = P0.move;
    const long base = (long)lb0 * TOPT_BLOCK * TOPT_OC_PER_THREAD + t;
#pragma unroll
    for (int kk = 0; kk < TOPT_OC_PER_THREAD; ++kk) {
      const long e = base + (long)kk * TOPT_BLOCK;
      const bool in = e < ne;
      const long ec = in ? e : 0;
      const double r = gp(P0.rho)[ec];
      const double g = -gp(P0.dcf)[ec];
      const double a = r * sqrt(g > 0.0 ? g : 0.0);
      const double rl = (r - mv) > 0.001 ? (r - mv) : 0.001;
      const double rh = (r + mv) < 1.0 ? (r + mv) : 1.0;
      ca[kk] = in ? a : 0.0;
      crl[kk] = in ? rl : 0.0;
      crh[kk] = in ? rh : 0.0;
    }
  }
  __syncthreads();
  for (int pass = 0; pass < TOPT_OCC_PASSES; ++pass) {
    bool any = false;
    if (!s_stop[0]) {
      any = true;
      occ_pass<D, true>(P0, lb0, 0, s_v0[0], pass, qs, red, vtot, rsh, rsf, s_lo, s_hi, s_mid, s_steps, s_stop, ca, crl, crh);
    }
    for (int k = 1; k < K; ++k) {
      if (s_stop[k]) continue;
      any = true;
      int lb;
      PD& P = occ_prob(tab, bmap, b + k * G, lb);
      occ_pass<D, false>(P, lb, k, s_v0[k], pass, qs, red, vtot, rsh, rsf, s_lo, s_hi, s_mid, s_steps, s_stop, ca, crl, crh);
    }
    if (!any) break;   // every held problem has stopped
    // the new bracket of every running held problem
    if (t == 0) {
      for (int k = 0; k < K; ++k) {
        if (s_stop[k]) continue;
        int lb;
        PD& P = occ_prob(tab, bmap, b + k * G, lb);
        OcState* o = P.oc;
        if (!occ_wait(&o->ver, s_v0[k] + (unsigned)pass + 1u, bar, lim)) {
          s_ok = 0;
          break;
        }
        s_lo[k] = occ_get(&o->lo);
        s_hi[k] = occ_get(&o->hi);
        s_mid[k] = occ_get(&o->last_mid);
        s_steps[k] = occ_get(&o->steps);
        s_stop[k] = occ_get(&o->stopped);
      }
    }
    __syncthreads();
    if (!s_ok) break;
  }
  // commit at the last evaluated lambda (k_oc_apply) and the statistics
  for (int k = 0; k < K; ++k) {
    int lb;
    PD& P = occ_prob(tab, bmap, b + k * G, lb);
    OcState* o = P.oc;
    if (!s_ok) continue;   // aborted: rho stays as it was
    const long ne = (long)P.nelx * P.nely;
    const double lam = s_mid[k], mv = P.move;
    double vsum = 0.0, vmax = 0.0;
    const long base = (long)lb * TOPT_BLOCK * TOPT_OC_PER_THREAD + t;
    for (int kk = 0; kk < TOPT_OC_PER_THREAD; ++kk) {
      const long e = base + (long)kk * TOPT_BLOCK;
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
    vsum = wave_sum(vsum);
    vmax = wave_max(vmax);
    if (lane == 0) {
      ared[wv] = vsum;
      ared[4 + wv] = vmax;
    }
    __syncthreads();
    double v2[2] = {0.0, 0.0};
    if (t == 0) {
      v2[0] = ared[0] + ared[1] + ared[2] + ared[3];
      double m = ared[4];
      for (int w = 5; w < 8; ++w) m = ared[w] > m ? ared[w] : m;
      v2[1] = m;
    }
    double tot[2];
    if (tree_reduce<2>(v2, P.opart, P.ctr[G_OC], lb, P.nb[G_OC], tot, 2u) && t == 0) {
      const double sv = tot[0], m = tot[1];
      o->vol = sv / (double)ne;
      o->change = m;
      o->iter += 1;
      o->done = ((o->iter > 5 && m < P.oc_tol) || o->iter >= P.oc_max_iter) ? 1 : 0;
    }
    __syncthreads();   // ared is reused by the next logical block
  }
}

}  // namespace topt
