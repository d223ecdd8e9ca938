// 3-D H8 kernels (trilinear hexahedra on the structured grid), FP64.
//
// Replaces the reference's 3-D Taichi kernels (paths relative to
// src/toptimization/):
//   fem/kernels.py:104-129        compute_diagonal_3d          -> k3_diag
//   fem/kernels.py:179-212        compute_matvec_3d            -> k3_matvec
//   fem/fused_kernels.py:59-94    matvec_and_dot_3d            -> k3_matvec (dot)
//   fem/kernels.py:275-308        compute_sensitivity_3d       -> k3_sens
//   fem/preconditioner.py:33-44,111-156  element-block Jacobi  -> k3_elemblock
//   optimizer/filter.py:75-97     3-D neighbour tables         -> k3_filter_tables
//   optimizer/filter.py:106-171   filters on those tables      -> k3_filter
//   fem/solver.py:27-167          PCG (device state machine)   -> k3_cg_*
//
// Layout: node n = (ix*(nely+1) + iy)*(nelz+1) + iz (z fastest), three DOFs
// per node interleaved (x, y, z); element e = (ex*nely + ey)*nelz + ez;
// fixed DOFs as a bitmask (bit 3n+c). Local nodes of an element follow
// mesh.py:91-100: the bottom face (z = ez) counter-clockwise from (0,0), then
// the top face. No edof table and no atomics: every kernel gathers from the
// 27-node (or 8-element) neighbourhood of its own node or element.
#pragma once
#include "topt_dev.h"

namespace topt {

// corner offsets of local node k (mesh.py:91-100) and the inverse map
__host__ __device__ constexpr int h8x(int k) { return (((k & 3) + 1) >> 1) & 1; }
__host__ __device__ constexpr int h8y(int k) { return ((k & 3) >> 1) & 1; }
__host__ __device__ constexpr int h8z(int k) { return k >> 2; }
__host__ __device__ constexpr int h8loc(int dx, int dy, int dz) {
  return (dy ? (dx ? 2 : 3) : (dx ? 1 : 0)) + 4 * dz;
}

struct Tap3 { int dx, dy, dz, pad; double w; };   // filter stencil entry, reference table order

// PCG scalars of a 3-D solve (device resident).
struct Cg3 {
  double rz, pAp, rr, alpha, beta, tol2, fnorm2, pad;
  int it, status, max_it, pad2;
};

struct Grid3 {
  int nelx, nely, nelz, precond;
  double Emin, penal, alpha_reg;
  const double* Ke;      // 24x24 row-major (material.py:89-171)
  const double* Ki;      // inv(Ke + alpha_reg I) (fields.py:139-143), element-block Jacobi
  const uint32_t* fix;   // bit 3n+c
  const double* s;       // element scales E_min + rho^p (1 - E_min)
  double* part;          // reduction partials
  unsigned* ctr;         // reduction counters (zero between launches)
  Cg3* cg;
  int nbn, nbe;          // node and element block counts
};

__device__ __forceinline__ unsigned fixbit3(const uint32_t* fix, long d) {
  return (gp(fix)[d >> 5] >> (unsigned)(d & 31)) & 1u;
}

__device__ __forceinline__ void node3(const Grid3& G, long n, int& ix, int& iy, int& iz) {
  const int nz1 = G.nelz + 1, ny1 = G.nely + 1;
  iz = (int)(n % nz1);
  const long t = n / nz1;
  iy = (int)(t % ny1);
  ix = (int)(t / ny1);
}

__device__ __forceinline__ long node_id3(const Grid3& G, int ix, int iy, int iz) {
  return ((long)ix * (G.nely + 1) + iy) * (G.nelz + 1) + iz;
}

// Elements sharing a node along one axis: 1 at the ends, 2 inside.
__device__ __forceinline__ int axis_count(int i, int nel) { return (i == 0 || i == nel) ? 1 : 2; }

// Values of a 3-vector field at the 27 nodes around (ix, iy, iz), times an
// optional per-node weight 1/sqrt(count) (0 outside the grid), and the
// scales of the 8 surrounding elements (0 outside).
template <bool WEIGHTED>
__device__ __forceinline__ void gather27(const Grid3& G, const double* v, int ix, int iy, int iz,
                                         double (&q)[27][3]) {
  const GLB double* V = gp(v);
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int jx = ix + i / 9 - 1, jy = iy + (i / 3) % 3 - 1, jz = iz + i % 3 - 1;
    const bool in = jx >= 0 && jx <= G.nelx && jy >= 0 && jy <= G.nely && jz >= 0 && jz <= G.nelz;
    const long m = in ? node_id3(G, jx, jy, jz) : 0;
    double w = in ? 1.0 : 0.0;
    if (WEIGHTED && in)
      w = 1.0 / sqrt((double)(axis_count(jx, G.nelx) * axis_count(jy, G.nely) * axis_count(jz, G.nelz)));
#pragma unroll
    for (int d = 0; d < 3; ++d) q[i][d] = in ? w * V[3 * m + d] : 0.0;
  }
}

__device__ __forceinline__ void scales8(const Grid3& G, const double* s, int ix, int iy, int iz, double (&se)[8],
                                        bool elemblock) {
  const GLB double* S = gp(s);
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ex = ix - 1 + (o >> 2), ey = iy - 1 + ((o >> 1) & 1), ez = iz - 1 + (o & 1);
    const bool in = ex >= 0 && ex < G.nelx && ey >= 0 && ey < G.nely && ez >= 0 && ez < G.nelz;
    double v = in ? S[((long)ex * G.nely + ey) * G.nelz + ez] : 0.0;
    if (elemblock && in) {
      const double sc = v > G.alpha_reg ? v : G.alpha_reg;   // preconditioner.py:140-141
      v = v / (sc * sc);
    }
    se[o] = v;
  }
}

// out[c] = sum over the 8 elements o of se[o] * sum_b,d M[3a+c][3b+d] q(b, d),
// a = our local node in element o. M read through the constant address space
// (uniform indices: scalar loads).
__device__ __forceinline__ void h8_gather(const double* Mp, const double (&se)[8], const double (&q)[27][3],
                                          double (&out)[3]) {
  const CST double* M = cp(Mp);
  out[0] = out[1] = out[2] = 0.0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ox = o >> 2, oy = (o >> 1) & 1, oz = o & 1;   // element (ix-1+ox, iy-1+oy, iz-1+oz)
    const int a = h8loc(1 - ox, 1 - oy, 1 - oz);
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int i = (ox + h8x(b)) * 9 + (oy + h8y(b)) * 3 + (oz + h8z(b));   // neighbour offset + 1
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[c] = fma(M[(3 * a + c) * 24 + 3 * b + d], q[i][d], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = fma(se[o], acc[c], out[c]);
  }
}

// s_e = E_min + rho_e^p (1 - E_min)  (kernels.py:200, 300)
__global__ __launch_bounds__(TOPT_BLOCK) void k3_scale(Grid3 G, const double* rho, double* s) {
  const long ne = (long)G.nelx * G.nely * G.nelz;
  const long e = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x;
  if (e >= ne) return;
  gp(s)[e] = G.Emin + pow(gp(rho)[e], G.penal) * (1.0 - G.Emin);
}

// diag K(s): sum_e s_e Ke_ii, fixed -> 1, clamp 1e-20 (kernels.py:104-129).
__global__ __launch_bounds__(TOPT_BLOCK) void k3_diag(Grid3 G, double* diag) {
  const long nn = (long)(G.nelx + 1) * (G.nely + 1) * (G.nelz + 1);
  const long n = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x;
  if (n >= nn) return;
  int ix, iy, iz;
  node3(G, n, ix, iy, iz);
  double se[8];
  scales8(G, G.s, ix, iy, iz, se, false);
  const CST double* K = cp(G.Ke);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double dsum = 0.0;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int a = h8loc(1 - (o >> 2), 1 - ((o >> 1) & 1), 1 - (o & 1));
      dsum = fma(se[o], K[(3 * a + c) * 25], dsum);
    }
    const long i = 3 * n + c;
    if (fixbit3(G.fix, i)) dsum = 1.0;
    else if (dsum < 1e-20) dsum = 1e-20;
    gp(diag)[i] = dsum;
  }
}

// PCG phases a matvec launch can run in (the state machine of solver.py:27-167).
enum Mv3 : int { MV3_PLAIN = 0, MV3_DOT = 1, MV3_CG = 2 };

// Ap = K(s) p, fixed rows identity (kernels.py:179-212). MV3_DOT: p.Ap into
// out[0] (fused_kernels.py:59-94). MV3_CG: skip unless the solve runs; the
// last block sets alpha = rz / pAp, or stops on |pAp| < 1e-30 (solver.py:114-117).
// Tiled: a workgroup covers MV3_TX x MV3_TY columns of MV3_TZ nodes (one wave
// per column, z fastest) and first stages the tile's node patch of p (with
// its one-node ring, zero outside the grid) and the scales of the elements
// around it in LDS with coalesced loads. A thread then takes its 8 elements
// in pairs along z: the 12 nodes a pair touches (36 values) come from LDS
// into registers, the pair's rows of Ke from scalar loads. The sums run in
// the order of h8_gather (elements o = 0..7, corners b, components d), so the
// result is that gather's, bit for bit.
#define MV3_TX 2
#define MV3_TY 2
#define MV3_TZ 64
static_assert(MV3_TX * MV3_TY * MV3_TZ == TOPT_BLOCK, "one node per thread");
__host__ __device__ inline int mv3_tiles(int nelx, int nely, int nelz) {
  return ((nelx + MV3_TX) / MV3_TX) * ((nely + MV3_TY) / MV3_TY) * ((nelz + MV3_TZ) / MV3_TZ);
}
__global__ __launch_bounds__(TOPT_BLOCK) void k3_matvec(Grid3 G, const double* p, double* Ap, int mode,
                                                        double* out) {
  constexpr int PX = MV3_TX + 2, PY = MV3_TY + 2, PZ = MV3_TZ + 2;   // node patch
  constexpr int EX = MV3_TX + 1, EY = MV3_TY + 1, EZ = MV3_TZ + 1;   // elements around it
  constexpr int NP = PX * PY * PZ, NE = EX * EY * EZ;
  __shared__ double sp[NP * 3];
  __shared__ double ss[NE];
  __shared__ double sh[4];
  if (mode == MV3_CG && G.cg->status != ST_RUN) return;
  const int nx1 = G.nelx + 1, ny1 = G.nely + 1, nz1 = G.nelz + 1;
  const int ntz = (nz1 + MV3_TZ - 1) / MV3_TZ, nty = (ny1 + MV3_TY - 1) / MV3_TY;
  const int b = (int)blockIdx.x;
  const int tz = b % ntz, t2 = b / ntz;
  const int ty = t2 % nty, tx = t2 / nty;
  const int x0 = tx * MV3_TX, y0 = ty * MV3_TY, z0 = tz * MV3_TZ;
  const int t = (int)threadIdx.x;
  const GLB double* P = gp(p);
  const GLB double* S = gp(G.s);
  {
    // every load of the thread issued before its LDS stores
    constexpr int KP = (NP + TOPT_BLOCK - 1) / TOPT_BLOCK, KE = (NE + TOPT_BLOCK - 1) / TOPT_BLOCK;
    double v[KP][3], e[KE];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int u = t + k * TOPT_BLOCK;
      const int px = u / (PY * PZ), py = (u / PZ) % PY, pz = u % PZ;
      const int jx = x0 - 1 + px, jy = y0 - 1 + py, jz = z0 - 1 + pz;
      const bool in = u < NP && jx >= 0 && jx < nx1 && jy >= 0 && jy < ny1 && jz >= 0 && jz < nz1;
      const long m = in ? ((long)jx * ny1 + jy) * nz1 + jz : 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) v[k][d] = in ? P[3 * m + d] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int u = t + k * TOPT_BLOCK;
      const int ex = x0 - 1 + u / (EY * EZ), ey = y0 - 1 + (u / EZ) % EY, ez = z0 - 1 + u % EZ;
      const bool in = u < NE && ex >= 0 && ex < G.nelx && ey >= 0 && ey < G.nely && ez >= 0 && ez < G.nelz;
      e[k] = in ? S[((long)ex * G.nely + ey) * G.nelz + ez] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int u = t + k * TOPT_BLOCK;
      if (u < NP) {
#pragma unroll
        for (int d = 0; d < 3; ++d) sp[3 * u + d] = v[k][d];
      }
    }
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int u = t + k * TOPT_BLOCK;
      if (u < NE) ss[u] = e[k];
    }
  }
  __syncthreads();
  const int w = t / MV3_TZ, cz = t % MV3_TZ;
  const int cx = w / MV3_TY, cy = w % MV3_TY;
  const int ix = x0 + cx, iy = y0 + cy, iz = z0 + cz;
  const bool act = ix < nx1 && iy < ny1 && iz < nz1;
  double dot[1] = {0.0};
  if (act) {
    const CST double* M = cp(G.Ke);
    double r[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int ox = 0; ox < 2; ++ox)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy) {
        // nodes (ix - 1 + ox + bx, iy - 1 + oy + by, iz - 1 + z), bx, by in {0, 1}, z in {0, 1, 2}:
        // patch position (cx + ox + bx, cy + oy + by, cz + z)
        double q[2][2][3][3];
#pragma unroll
        for (int bx = 0; bx < 2; ++bx)
#pragma unroll
          for (int by = 0; by < 2; ++by)
#pragma unroll
            for (int z = 0; z < 3; ++z) {
              const int u = ((cx + ox + bx) * PY + (cy + oy + by)) * PZ + (cz + z);
#pragma unroll
              for (int d = 0; d < 3; ++d) q[bx][by][z][d] = sp[3 * u + d];
            }
#pragma unroll
        for (int oz = 0; oz < 2; ++oz) {
          // element (ix - 1 + ox, iy - 1 + oy, iz - 1 + oz); our local node a
          const int a = h8loc(1 - ox, 1 - oy, 1 - oz);
          const double se = ss[((cx + ox) * EY + (cy + oy)) * EZ + (cz + oz)];
          double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
          for (int bb = 0; bb < 8; ++bb) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int d = 0; d < 3; ++d)
                acc[c] = fma(M[(3 * a + c) * 24 + 3 * bb + d], q[h8x(bb)][h8y(bb)][oz + h8z(bb)][d], acc[c]);
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) r[c] = fma(se, acc[c], r[c]);
        }
      }
    const long n = ((long)ix * ny1 + iy) * nz1 + iz;
    const int uc = ((cx + 1) * PY + (cy + 1)) * PZ + (cz + 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long i = 3 * n + c;
      const double pc = sp[3 * uc + c];
      const double v = fixbit3(G.fix, i) ? pc : r[c];
      gp(Ap)[i] = v;
      dot[0] = fma(pc, v, dot[0]);
    }
  }
  if (mode == MV3_PLAIN) return;
  double tot[1];
  if (!grid_sum<1>(dot, G.part, G.ctr, blockIdx.x, gridDim.x, tot, sh)) return;
  if (threadIdx.x != 0) return;
  if (mode == MV3_DOT) {
    out[0] = tot[0];
    return;
  }
  Cg3* st = G.cg;
  st->pAp = tot[0];
  if (fabs(tot[0]) < 1e-30) {
    st->status = ST_BREAK;
    st->it = st->max_it;          // reported as max_iter (solver.py:157-159)
  } else {
    st->alpha = st->rz / tot[0];
  }
}

// Element-block Jacobi z = C^-1/2 sum_e w_e P_e^T Ki P_e C^-1/2 r, fixed -> 0,
// with r.z and r.r (preconditioner.py:111-156); w_e = s_e / max(s_e, alpha)^2.
// mode 0: op-level, (r.z, r.r) into out[0..1]; 1: PCG setup (also p = z;
// the last block sets rz); 2: PCG iteration (the last block finishes the
// step: convergence, beta). Skips unless the solve runs (modes 1, 2).
__global__ __launch_bounds__(TOPT_BLOCK) void k3_elemblock(Grid3 G, const double* r, double* z, double* p, int mode,
                                                           double* out);

// The last block's end of a PCG iteration (solver.py:141-152): it += 1,
// convergence on r.r, the iteration cap, beta and rz.
__device__ __forceinline__ void cg3_finish(Cg3* st, double rz_new, double rr) {
  st->it += 1;
  st->rr = rr;
  if (rr < st->tol2) {
    st->status = ST_CONV;
  } else if (st->it >= st->max_it) {
    st->status = ST_MAXIT;
  } else {
    st->beta = rz_new / (st->rz + 1e-60);
    st->rz = rz_new;
  }
}

__global__ __launch_bounds__(TOPT_BLOCK) void k3_elemblock(Grid3 G, const double* r, double* z, double* p, int mode,
                                                           double* out) {
  __shared__ double sh[8];
  if (mode != 0 && G.cg->status != ST_RUN) return;
  const long nn = (long)(G.nelx + 1) * (G.nely + 1) * (G.nelz + 1);
  const long n = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x;
  double v2[2] = {0.0, 0.0};
  if (n < nn) {
    int ix, iy, iz;
    node3(G, n, ix, iy, iz);
    double q[27][3], se[8], acc[3];
    gather27<true>(G, r, ix, iy, iz, q);
    scales8(G, G.s, ix, iy, iz, se, true);
    h8_gather(G.Ki, se, q, acc);
    const double wn =
        1.0 / sqrt((double)(axis_count(ix, G.nelx) * axis_count(iy, G.nely) * axis_count(iz, G.nelz)));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long i = 3 * n + c;
      const double zc = fixbit3(G.fix, i) ? 0.0 : acc[c] * wn;
      const double rc = gp(r)[i];
      gp(z)[i] = zc;
      if (mode == 1) gp(p)[i] = zc;
      v2[0] = fma(rc, zc, v2[0]);
      v2[1] = fma(rc, rc, v2[1]);
    }
  }
  double tot[2];
  if (!grid_sum<2>(v2, G.part, G.ctr, blockIdx.x, gridDim.x, tot, sh)) return;
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    out[0] = tot[0];
    out[1] = tot[1];
  } else if (mode == 1) {
    G.cg->rz = tot[0];
  } else {
    cg3_finish(G.cg, tot[0], tot[1]);
  }
}

// PCG setup (solver.py:55-104): with Ap = K u already formed from the
// (zeroed-at-fixed) start u: r = f - Ap; Jacobi z = r / diag, none z = r,
// p = z (element-block: z and p come from k3_elemblock mode 1 afterwards);
// the last block sets |f|^2, the tolerance, rz and the status (zero load:
// u = 0, 0 iterations).
__global__ __launch_bounds__(TOPT_BLOCK) void k3_cg_init(Grid3 G, const double* f, const double* Ap, double* u,
                                                         double* r, double* z, double* p, const double* diag,
                                                         double tol, int max_it) {
  __shared__ double sh[8];
  const long nd = 3 * (long)(G.nelx + 1) * (G.nely + 1) * (G.nelz + 1);
  double v2[2] = {0.0, 0.0};
  const long i0 = (long)blockIdx.x * (3 * TOPT_BLOCK) + threadIdx.x;   // as in k3_cg_update
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long i = i0 + c * TOPT_BLOCK;
    if (i >= nd) break;
    const double fi = gp(f)[i];
    const double ri = fi - gp(Ap)[i];
    gp(r)[i] = ri;
    if (G.precond != PC_ELEMBLOCK) {
      const double zi = G.precond == PC_JACOBI ? ri / gp(diag)[i] : ri;
      gp(z)[i] = zi;
      gp(p)[i] = zi;
      v2[0] = fma(ri, zi, v2[0]);
    }
    v2[1] = fma(fi, fi, v2[1]);
  }
  double tot[2];
  if (!grid_sum<2>(v2, G.part, G.ctr, blockIdx.x, gridDim.x, tot, sh)) return;
  if (threadIdx.x != 0) return;
  Cg3* st = G.cg;
  st->fnorm2 = tot[1];
  st->tol2 = (tol * sqrt(tot[1])) * (tol * sqrt(tot[1]));
  st->rz = tot[0];
  st->it = 0;
  st->max_it = max_it;
  st->beta = 0.0;
  st->status = sqrt(tot[1]) < 1e-30 ? ST_ZERO : (max_it > 0 ? ST_RUN : ST_MAXIT);
}

// u = 0 after a zero-load setup (solver.py:60-63).
__global__ __launch_bounds__(TOPT_BLOCK) void k3_cg_zero_u(Grid3 G, double* u, long nd) {
  if (G.cg->status != ST_ZERO) return;
  for (long i = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x; i < nd; i += (long)gridDim.x * TOPT_BLOCK)
    gp(u)[i] = 0.0;
}

// u += alpha p; r -= alpha Ap; Jacobi z = r/diag, none z = r, then r.z and
// r.r and the end of the iteration (fused_kernels.py:110-166); element-block
// Jacobi only updates u and r here (fused_kernels.py:169-180).
__global__ __launch_bounds__(TOPT_BLOCK) void k3_cg_update(Grid3 G, double* u, double* r, double* z, const double* p,
                                                           const double* Ap, const double* diag, long nd) {
  __shared__ double sh[8];
  const Cg3* st = G.cg;
  if (st->status != ST_RUN) return;
  const double alpha = st->alpha;
  double v2[2] = {0.0, 0.0};
  // a workgroup takes 3 x TOPT_BLOCK consecutive DOFs, a thread every
  // TOPT_BLOCK-th of them (each load instruction of a wave contiguous)
  const long i0 = (long)blockIdx.x * (3 * TOPT_BLOCK) + threadIdx.x;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long i = i0 + c * TOPT_BLOCK;
    if (i >= nd) break;
    gp(u)[i] = fma(alpha, gp(p)[i], gp(u)[i]);
    const double ri = fma(-alpha, gp(Ap)[i], gp(r)[i]);
    gp(r)[i] = ri;
    if (G.precond != PC_ELEMBLOCK) {
      const double zi = G.precond == PC_JACOBI ? ri / gp(diag)[i] : ri;
      gp(z)[i] = zi;
      v2[0] = fma(ri, zi, v2[0]);
      v2[1] = fma(ri, ri, v2[1]);
    }
  }
  if (G.precond == PC_ELEMBLOCK) return;
  double tot[2];
  if (!grid_sum<2>(v2, G.part, G.ctr, blockIdx.x, gridDim.x, tot, sh)) return;
  if (threadIdx.x == 0) cg3_finish(G.cg, tot[0], tot[1]);
}

// p = z + beta p (fused_kernels.py:183-187).
__global__ __launch_bounds__(TOPT_BLOCK) void k3_cg_pupdate(Grid3 G, const double* z, double* p, long nd) {
  const Cg3* st = G.cg;
  if (st->status != ST_RUN) return;
  const double beta = st->beta;
  const long i0 = (long)blockIdx.x * (3 * TOPT_BLOCK) + threadIdx.x;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long i = i0 + c * TOPT_BLOCK;
    if (i >= nd) break;
    gp(p)[i] = fma(beta, gp(p)[i], gp(z)[i]);
  }
}

// Compliance and sensitivities (kernels.py:275-308): per element u_e^T Ke u_e,
// dc = -p rho^(p-1) (1 - E_min) u^T Ke u, c = sum s_e u^T Ke u into out[0].
__global__ __launch_bounds__(TOPT_BLOCK) void k3_sens(Grid3 G, const double* u, const double* rho, double* dc,
                                                      double* out) {
  __shared__ double sh[4];
  const long ne = (long)G.nelx * G.nely * G.nelz;
  const long e = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x;
  double c1[1] = {0.0};
  if (e < ne) {
    const int ez = (int)(e % G.nelz);
    const long t = e / G.nelz;
    const int ey = (int)(t % G.nely), ex = (int)(t / G.nely);
    double ue[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const long m = node_id3(G, ex + h8x(b), ey + h8y(b), ez + h8z(b));
#pragma unroll
      for (int d = 0; d < 3; ++d) ue[3 * b + d] = gp(u)[3 * m + d];
    }
    const CST double* K = cp(G.Ke);
    double uKu = 0.0;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
      double kr = 0.0;
#pragma unroll
      for (int j = 0; j < 24; ++j) kr = fma(K[i * 24 + j], ue[j], kr);
      uKu = fma(ue[i], kr, uKu);
    }
    const double re = gp(rho)[e];
    const double sv = G.Emin + pow(re, G.penal) * (1.0 - G.Emin);
    gp(dc)[e] = -G.penal * pow(re, G.penal - 1.0) * (1.0 - G.Emin) * uKu;
    c1[0] = sv * uKu;
  }
  double tot[1];
  if (!grid_sum<1>(c1, G.part, G.ctr, blockIdx.x, gridDim.x, tot, sh)) return;
  if (threadIdx.x == 0) out[0] = tot[0];
}

// Filters over the stencil in the reference's table order (filter.py:85-94:
// jx outer, jy, jz inner), FP64 sums; modes as the 2-D k_filter:
//   F_HS: Hs = sum w;  F_DENSITY: rhof = sum w rho / (Hs + 1e-20);
//   F_DCAVG: sum w dc / (Hs + 1e-20);  F_SENS: sum w rho dc / (rho_e Hs + 1e-20).
__global__ __launch_bounds__(TOPT_BLOCK) void k3_filter(int nelx, int nely, int nelz, const Tap3* taps, int ntaps,
                                                        int mode, const double* rho, const double* dc,
                                                        const double* Hs, double* out) {
  const long ne = (long)nelx * nely * nelz;
  const long e = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x;
  if (e >= ne) return;
  const int ez = (int)(e % nelz);
  const long t = e / nelz;
  const int ey = (int)(t % nely), ex = (int)(t / nely);
  const CST Tap3* tp = cp(taps);
  double num = 0.0;
  for (int k = 0; k < ntaps; ++k) {
    const int jx = ex + tp[k].dx, jy = ey + tp[k].dy, jz = ez + tp[k].dz;
    if (jx < 0 || jx >= nelx || jy < 0 || jy >= nely || jz < 0 || jz >= nelz) continue;
    const long j = ((long)jx * nely + jy) * nelz + jz;
    double v;
    switch (mode) {
      case F_HS: v = 1.0; break;
      case F_DENSITY: v = gp(rho)[j]; break;
      case F_DCAVG: v = gp(dc)[j]; break;
      default: v = gp(rho)[j] * gp(dc)[j]; break;
    }
    num = fma(tp[k].w, v, num);
  }
  switch (mode) {
    case F_HS: gp(out)[e] = num; break;
    case F_SENS: gp(out)[e] = num / (gp(rho)[e] * gp(Hs)[e] + 1e-20); break;
    default: gp(out)[e] = num / (gp(Hs)[e] + 1e-20); break;
  }
}

// The reference's padded 3-D neighbour tables (filter.py:75-97): neighbours
// (-1 pad), float32 weights (0 pad), counts; one thread per element.
__global__ __launch_bounds__(TOPT_BLOCK) void k3_filter_tables(int nelx, int nely, int nelz, const Tap3* taps,
                                                               int ntaps, int width, int* nb, float* wt, int* cnt) {
  const long ne = (long)nelx * nely * nelz;
  const long e = (long)blockIdx.x * TOPT_BLOCK + threadIdx.x;
  if (e >= ne) return;
  const int ez = (int)(e % nelz);
  const long t = e / nelz;
  const int ey = (int)(t % nely), ex = (int)(t / nely);
  int k = 0;
  for (int q = 0; q < ntaps; ++q) {
    const int jx = ex + taps[q].dx, jy = ey + taps[q].dy, jz = ez + taps[q].dz;
    if (jx < 0 || jx >= nelx || jy < 0 || jy >= nely || jz < 0 || jz >= nelz) continue;
    nb[e * width + k] = (int)(((long)jx * nely + jy) * nelz + jz);
    wt[e * width + k] = (float)taps[q].w;
    ++k;
  }
  for (int j = k; j < width; ++j) {
    nb[e * width + j] = -1;
    wt[e * width + j] = 0.0f;
  }
  cnt[e] = k;
}

}  // namespace topt
