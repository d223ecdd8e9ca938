// Device-side building blocks shared by the SIMP kernels (gfx950 / CDNA4).
//
// Layout in HBM (per design problem):
//   nodes  n = ix*(nely+1) + iy, DOF vectors interleaved per node (x, y)
//          -> one float2/double2 per node;
//   elements e = ex*nely + ey (one scalar per element);
//   fixed DOFs: bitmask, bit (2n + c) of a uint32 array.
// (reference numbering: src/toptimization/mesh.py:3-24, problem.py:167-173)
//
// A batch is a table of ProbDev descriptors; every kernel runs over a
// "node grid" or an "element grid" whose blocks are split between problems
// (blk2prob map). A single problem is a batch of one, passed by value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TOPT_BLOCK 256
#define TOPT_WAVE 64

namespace topt {

// ---------------------------------------------------------------- constants
enum Precond : int { PC_NONE = 0, PC_JACOBI = 1, PC_ELEMBLOCK = 2, PC_BLOCK2 = 3 };
enum Status : int { ST_RUN = 0, ST_CONV = 1, ST_BREAK = 2, ST_MAXIT = 3, ST_ZERO = 4 };
enum Mode : int { MODE_CG = 0, MODE_REFINE = 1 };
enum FilterMode : int { F_DENSITY = 0, F_DCAVG = 1, F_SENS = 2, F_TRANSPOSE = 3, F_HS = 4 };

struct Tap { int dx, dy; double w; };  // filter stencil entry (reference table order)

// PCG scalars of one problem (device resident, FP64).
struct CgState {
  double rz, rz_prev, alpha, beta, pAp, rr, rr_true, rmax2;
  int it, status, mode, pending;
  int pcur, first, refines, pad;
  unsigned cnt[4];          // arrival counters for last-block reductions
};

// OC / SIMP per-iteration scalars of one problem.
struct OcState {
  double lo, hi, last_mid, c, vol, change;
  int steps, stopped, iter, pad;
  unsigned cnt[4];
};

// Block grids a batch is launched over; each problem owns nb[g] consecutive
// logical blocks starting at b0[g].
enum Grid : int { G_NODE = 0, G_ELEM = 1, G_FILT = 2, G_OC = 3 };
#define TOPT_FILT_TY 64   // filter tile: 64 elements along y ...
#define TOPT_FILT_TX 8    // ... by 8 along x, 2 per thread
#define TOPT_OC_PER_THREAD 8
#define TOPT_OC_MAXD 6    // bisection levels evaluated per OC pass (<= 6)

struct ProbDev {
  int nelx, nely;
  int nb[4], b0[4];
  int precond, max_it, ntaps, R;
  double Emin, penal, tol2, delta2;   // tol2 = (tol |f|)^2; delta2 = reliable-update ratio^2
  double alpha_reg, volfrac, move, fnorm2;
  float Ke32[64];
  double Ke64[64];
  float Ki32[64];        // regularised Ke inverse (element-block Jacobi)
  double Ki64[64];
  // FEM vectors (node interleaved). T = float (mixed) or double (plain).
  void* x;   // CG iterate: plain = the solution; mixed = FP32 correction
  void* r; void* z; void* p0; void* p1; void* Ap;
  void* pre; void* pre2;   // preconditioner data
  void* s;                 // element scale in T
  double* s64;             // element scale FP64 (mixed: residual refresh)
  double* u64;             // FP64 solution (plain: == x)
  const double* f;         // FP64 load vector
  const uint32_t* fix;     // fixed-DOF bitmask
  // design fields (FP64, per element)
  double* rho; double* rhof; double* dc; double* dcf; const double* dv; double* Hs;
  const Tap* taps;
  CgState* cg; double* part;   // part: nb[G_NODE] * 2 doubles
  OcState* oc; double* opart;  // opart: max(nb[G_ELEM], nb[G_OC] * 64) doubles
};

// ------------------------------------------------------------ vector types
template <typename T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };
template <typename T> using vec2 = typename V2<T>::t;

template <typename T> __device__ __forceinline__ const T* keptr(const ProbDev& P);
template <> __device__ __forceinline__ const float* keptr<float>(const ProbDev& P) { return P.Ke32; }
template <> __device__ __forceinline__ const double* keptr<double>(const ProbDev& P) { return P.Ke64; }
template <typename T> __device__ __forceinline__ const T* kiptr(const ProbDev& P);
template <> __device__ __forceinline__ const float* kiptr<float>(const ProbDev& P) { return P.Ki32; }
template <> __device__ __forceinline__ const double* kiptr<double>(const ProbDev& P) { return P.Ki64; }

// ------------------------------------------------------------ block mapping
// The hardware hands workgroup b to XCD (b % 8). Remap so each XCD works on a
// contiguous range of logical blocks: its L2 then holds the neighbouring
// node columns the stencil re-reads.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int X = 8;
  const int q = n / X, r = n % X;
  const int x = b % X, j = b / X;
  return x * q + (x < r ? x : r) + j;
}

// Problem descriptor + local block index for this workgroup.
template <bool BATCHED>
__device__ __forceinline__ const ProbDev& get_prob(const ProbDev* tab, const int* bmap,
                                                   const ProbDev& single, int b) {
  if constexpr (BATCHED) return tab[bmap[b]];
  else return single;
}

__device__ __forceinline__ unsigned fixbits(const uint32_t* fix, long n) {
  return (fix[n >> 4] >> (2 * (unsigned)(n & 15))) & 3u;
}

// ------------------------------------------------------------ reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = TOPT_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, TOPT_WAVE);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = TOPT_WAVE / 2; o > 0; o >>= 1) { T w = __shfl_xor(v, o, TOPT_WAVE); v = v > w ? v : w; }
  return v;
}

// Sum NV values over the block; result valid in thread 0. Fixed order.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh /* >= 4*NV */) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[wv * NV + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = sh[k];
      for (int w = 1; w < nw; ++w) t += sh[w * NV + k];
      v[k] = t;
    }
  }
  __syncthreads();
}

// Last-block pattern: every block deposits NV partials (thread 0 holds them),
// the block that arrives last sums all partials of its problem in a fixed
// order (independent of arrival order and of other problems in the batch) and
// returns true with the totals in tot[] (valid in every thread).
template <int NV>
__device__ bool last_block_sum(const double (&v)[NV], double* part, unsigned* counter, int lb,
                               int nb, double (&tot)[NV], double* sh) {
  __shared__ int s_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      __hip_atomic_store(&part[(long)lb * NV + k], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    s_last = (t == (unsigned)(nb - 1));
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      acc[k] += __hip_atomic_load(&part[(long)j * NV + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<NV>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k] = acc[k];
    *counter = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = sh[k];
  __syncthreads();
  return true;
}

// ------------------------------------------------------------ Q4 stencil
// Node-centric gather of K(s) q at node (ix, iy) from the 3x3 neighbourhood
// values q[dx][dy] (dx, dy in 0..2 <-> offsets -1..+1) and the scales of the
// four adjacent elements sc[a][b] = s(ix-1+a, iy-1+b) (0 outside the grid).
// Element (ex, ey) has local nodes BL(0,0) BR(1,0) TR(1,1) TL(0,1); our node
// is its local node (1-a, 1-b).
template <typename T>
__device__ __forceinline__ void q4_gather(const T* __restrict__ K, const T (&sc)[2][2],
                                          const vec2<T> (&q)[3][3], T& ox, T& oy) {
  constexpr int CX[4] = {0, 1, 1, 0};
  constexpr int CY[4] = {0, 0, 1, 1};
  T ax = T(0), ay = T(0);
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      // local index of our node inside this element
      const int me = (a == 1) ? ((b == 1) ? 0 : 3) : ((b == 1) ? 1 : 2);
      T tx = T(0), ty = T(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const vec2<T> v = q[a + CX[k]][b + CY[k]];
        tx += K[(2 * me) * 8 + 2 * k] * v.x + K[(2 * me) * 8 + 2 * k + 1] * v.y;
        ty += K[(2 * me + 1) * 8 + 2 * k] * v.x + K[(2 * me + 1) * 8 + 2 * k + 1] * v.y;
      }
      ax += sc[a][b] * tx;
      ay += sc[a][b] * ty;
    }
  }
  ox = ax;
  oy = ay;
}

// Scales of the four elements around node (ix, iy); zero outside the grid.
template <typename T>
__device__ __forceinline__ void load_scales(const T* __restrict__ s, int ix, int iy, int nelx,
                                            int nely, T (&sc)[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int ex = ix - 1 + a, ey = iy - 1 + b;
      const bool ok = (ex >= 0) & (ex < nelx) & (ey >= 0) & (ey < nely);
      const int cx = ex < 0 ? 0 : (ex >= nelx ? nelx - 1 : ex);
      const int cy = ey < 0 ? 0 : (ey >= nely ? nely - 1 : ey);
      const T v = s[(long)cx * nely + cy];
      sc[a][b] = ok ? v : T(0);
    }
  }
}

// Clamped neighbour node index (values outside the grid only meet zero scales).
__device__ __forceinline__ long nbr_node(int ix, int iy, int dx, int dy, int nelx, int nely) {
  int jx = ix + dx, jy = iy + dy;
  jx = jx < 0 ? 0 : (jx > nelx ? nelx : jx);
  jy = jy < 0 ? 0 : (jy > nely ? nely : jy);
  return (long)jx * (nely + 1) + jy;
}

}  // namespace topt
