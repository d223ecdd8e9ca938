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
enum Precond : int { PC_NONE = 0, PC_JACOBI = 1, PC_ELEMBLOCK = 2, PC_BLOCK2 = 3, PC_MG = 4 };
enum Status : int { ST_RUN = 0, ST_CONV = 1, ST_BREAK = 2, ST_MAXIT = 3, ST_ZERO = 4, ST_DONE = 5 };
enum Mode : int { MODE_CG = 0, MODE_REFINE = 1 };
enum FilterMode : int { F_DENSITY = 0, F_DCAVG = 1, F_SENS = 2, F_TRANSPOSE = 3, F_HS = 4 };

struct Tap { int dx, dy; double w; };  // filter stencil entry (reference table order)

// PCG scalars of one problem (device resident, FP64).
// Double-buffered: the PCG slot kernel with sequence number q reads
// cg[q & 1] and its block 0 writes the advanced state to cg[(q + 1) & 1].
struct CgState {
  double rz, rz_prev, alpha, beta, pAp, rr, rr_true, rmax2;
  int it, status, op, pending;   // op: what the kernel that wrote this state executed
  int pad3, first, refines, amode;
  int cflag, upend, pad1, pad2;   // upend: u += alpha p of the last B_CG not applied yet
};
// PCG slot operations (the state machine of fem/solver.py:27-159 on the device).
enum SlotOp : int { OP_START = 0, OP_A_CG = 1, OP_A_REF = 2, OP_B_CG = 3, OP_B_REF = 4, OP_C = 5 };
enum CFlag : int { CF_CG = 0, CF_REFRESH = 1 };

// OC / SIMP per-iteration scalars of one problem.
struct OcState {
  double lo, hi, last_mid, c, vol, change;
  int steps, stopped, iter, done;   // done: the OC stopping rule has been met
  unsigned ver;                     // bracket version of the one-launch OC (k_oc_coop)
  unsigned pad;
};

// Block grids a batch is launched over; each problem owns nb[g] consecutive
// logical blocks starting at b0[g].
enum Grid : int { G_NODE = 0, G_ELEM = 1, G_FILT = 2, G_OC = 3, G_PCG = 4 };
#define TOPT_NGRID 5
#define TOPT_FILT_TY 64   // filter tile: 64 elements along y ...
#define TOPT_FILT_TX 8    // ... by 8 along x, 2 per thread
#ifndef TOPT_OC_PER_THREAD
#define TOPT_OC_PER_THREAD 8
#endif
#define TOPT_OC_MAXD 6    // bisection levels evaluated per OC pass (<= 6)
// nodes per thread in the PCG slot kernels (FP32 / FP64 vectors). Two in
// FP32 as in FP64: 8-column tiles give small grids twice the workgroups (the
// 300x100 MBB runs 76 per slot launch instead of 38, 8 % faster) and leave
// the 1200x400 bridge within 1 %.
#ifndef TOPT_NPT_F32
#define TOPT_NPT_F32 2
#endif
template <typename T> struct Npt { static constexpr int v = sizeof(T) == 4 ? TOPT_NPT_F32 : 2; };
// nodes per thread in the op-level matvec (FP32 / FP64): at 8192x4096 the
// 16-column FP32 tiles re-read less halo than 8-column ones
#ifndef TOPT_NPT_MV_F32
#define TOPT_NPT_MV_F32 4
#endif
#ifndef TOPT_NPT_MV_F64
#define TOPT_NPT_MV_F64 2
#endif
template <typename T> struct NptMv { static constexpr int v = sizeof(T) == 4 ? TOPT_NPT_MV_F32 : TOPT_NPT_MV_F64; };
// modal op-level matvec: walks of TOPT_MV_WALK node columns (3, 7, 15, 31
// or 63; short walks keep more loads in flight and no tail of long-running
// workgroups: at 8192x4096, 7 columns ran 124 us, 15 131 us, 63 146 us)
#ifndef TOPT_MV_WALK
#define TOPT_MV_WALK 7
#endif
#define TOPT_MV_LMIN 3

// One level of the geometric multigrid hierarchy (PC_MG). Level 0 is the
// design grid; level l+1 has ceil(nx/2) x ceil(ny/2) elements, each the
// agglomerate of a 2x2 patch of level-l elements (those past the level-l
// grid count as void). Node vectors hold 2 T per node, as the PCG vectors.
#define TOPT_MG_MAXLEV 14
// Node record of a sweep level (l >= 1): entry 18 d + 2 q + c is the coefficient
// of component c at neighbour q = 3 (dx + 1) + (dy + 1) in the node's row
// for component d; entries 36, 37 are 1/diag; 38, 39 are zero (16-byte
// multiple).
#define TOPT_MG_ST 40
struct MgLevel {
  int nx, ny;     // elements
  void* K;        // levels >= 1: Galerkin element matrices, 36 T each (packed upper
                  // triangle), entry-major per element column (mg_kel)
  void* x;        // coarsest level: the coarse solve's result (null elsewhere)
  void* b;        // right-hand side (level 0: the PCG residual r)
  void* r;        // levels 1 .. L-1: the corrected, post-smoothed iterate (null elsewhere)
  void* dinv;     // 1 / diag (level 0: the Jacobi data in pre); 0 on decoupled DOFs
  void* st;       // levels 1 .. L-1: one record of TOPT_MG_ST T per node (its
                  // two rows of the assembled level operator, then 1/diag);
                  // null elsewhere
};

struct ProbDev {
  int nelx, nely;
  int nb[TOPT_NGRID], b0[TOPT_NGRID];
  int precond, max_it, ntaps, R;
  double Emin, penal, tol2, delta2;   // tol2 = (tol |f|)^2; delta2 = reliable-update ratio^2
  double alpha_reg, volfrac, move, fnorm2;
  float Ke32[64];
  double Ke64[64];
  float Ki32[64];        // regularised Ke inverse (element-block Jacobi)
  double Ki64[64];
  double kc[6];          // Ke = KS pattern x kc when kstruct (see q4_gather_s)
  int kstruct, oc_max_iter;   // oc_max_iter, oc_tol: the OC stopping rule
  double oc_tol;
  double km[6];          // Ke = G^T W G with W from km when kmodal (see modal_weights)
  int kmodal, km_pad;
  // FEM vectors (node interleaved). T = float (mixed) or double (plain).
  void* x;   // CG iterate: aliases the FP64 solution u64 (both paths)
  void* r; void* z; void* p0; void* p1; void* Ap;
  // the residual's second buffer: plans whose problems all use multigrid
  // alternate r between r (parity 0) and r1 (parity 1) once per CG
  // iteration (the fused slot B + level-0 down sweep reads one and writes
  // the other); every other plan sets r1 = r
  void* r1;
  void* pre; void* pre2;   // preconditioner data
  void* s;                 // element scale in T
  double* s64;             // element scale FP64 (mixed: residual refresh)
  double* u64;             // FP64 solution (plain: == x)
  const double* f;         // FP64 load vector
  const uint32_t* fix;     // fixed-DOF bitmask
  // design fields (FP64, per element)
  double* rho; double* rhof; double* dc; double* dcf; const double* dv; double* Hs;
  const Tap* taps;
  CgState* cg; double* part;   // cg[2]; partials: PCG slots 2 x nb[G_PCG] x 2, node-grid tree
  OcState* oc; double* opart;  // elem / OC-grid tree partials
  unsigned* ctr[4];            // tree counters per grid (G_NODE, G_ELEM, -, G_OC)
  // multigrid preconditioner (PC_MG): levels 0..mg_nlev-1, the last one
  // solved with the dense inverse mg_inv (mg_ncd x mg_ncd, in T: FP32 on
  // the mixed path, FP64 on the plain one)
  int mg_nlev, mg_ncd;
  int mg_lt, mg_pad;          // first level of the single-workgroup tail
  double mg_omega;            // Jacobi damping of the smoother (set from Ke)
  void* mg_inv;
  MgLevel mg[TOPT_MG_MAXLEV];
};

// ------------------------------------------------------------ vector types
// Two-component node vectors as native clang vectors (not HIP's class
// types): assignable through address-space-qualified pointers and loaded
// as one dwordx2 / dwordx4.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
template <typename T> struct V2;
template <> struct V2<float> { using t = f2v; };
template <> struct V2<double> { using t = d2v; };
template <typename T> using vec2 = typename V2<T>::t;


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

// Descriptors are read through the constant address space: the table is
// never written by a kernel, so every field comes in with scalar loads
// (one per wave) instead of per-lane vector loads.
typedef const __attribute__((address_space(4))) ProbDev PD;
// Data pointers taken from a descriptor are generic; casting them to the
// global address space gives global_load/store instead of flat_load/store.
#define GLB __attribute__((address_space(1)))
template <typename X>
__device__ __forceinline__ GLB X* gp(X* p) { return (GLB X*)p; }
template <typename X>
__device__ __forceinline__ const GLB X* gp(const X* p) { return (const GLB X*)p; }

// Read-only tables (filter taps) through the constant address space: scalar loads.
#define CST __attribute__((address_space(4)))
template <typename X>
__device__ __forceinline__ const CST X* cp(const X* p) { return (const CST X*)p; }

// Problem descriptor of this workgroup (table entry or the by-value single).
template <bool BATCHED>
__device__ __forceinline__ PD& get_prob(const ProbDev* tab, const int* bmap,
                                        const ProbDev& single, int b) {
  if constexpr (BATCHED) return *(PD*)(tab + bmap[b]);
  else return *(PD*)&single;
}

// Batched grids: the block map holds (problem, local block) pairs, so a
// launch can cover any subset of a plan's problems (each with all its
// blocks); a null map stands for the identity map of a one-problem plan
// (problem 0, block b), which saves the map's load on the way to the
// descriptor. Single-problem launches pass the descriptor by value.
template <bool BATCHED>
__device__ __forceinline__ PD& grid_prob(const ProbDev* tab, const int2* bmap, const ProbDev& single, int b,
                                         int& lb) {
  if constexpr (BATCHED) {
    if (bmap == nullptr) {
      lb = b;
      return *(PD*)tab;
    }
    const int2 e = bmap[b];
    lb = e.y;
    return *(PD*)(tab + e.x);
  } else {
    lb = b;
    return *(PD*)&single;
  }
}

__device__ __forceinline__ unsigned fixbits(const uint32_t* fix, long n) {
  return (gp(fix)[n >> 4] >> (2 * (unsigned)(n & 15))) & 3u;
}

template <typename T> struct KeSel;
template <> struct KeSel<float> {
  static __device__ __forceinline__ const __attribute__((address_space(4))) float* ke(PD& P) { return P.Ke32; }
  static __device__ __forceinline__ const __attribute__((address_space(4))) float* ki(PD& P) { return P.Ki32; }
};
template <> struct KeSel<double> {
  static __device__ __forceinline__ const __attribute__((address_space(4))) double* ke(PD& P) { return P.Ke64; }
  static __device__ __forceinline__ const __attribute__((address_space(4))) double* ki(PD& P) { return P.Ki64; }
};

// ------------------------------------------------------------ reductions
// Wavefront reductions of doubles with DPP lane moves (VALU, no LDS
// round trip): pairs (quad_perm 1,0,3,2), quads (2,3,0,1), half rows
// (row_half_mirror), rows (row_mirror), then row_bcast:15 into rows 1 and 3
// and row_bcast:31 into rows 2 and 3, which leaves the total in lane 63.
// Fixed combination order: bitwise reproducible. Every lane must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v, double ident) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(ident);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// Sum over the wavefront, returned to every lane.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1, 0xF>(v, 0.0);    // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E, 0xF>(v, 0.0);    // quad_perm [2,3,0,1]
  v += dpp_f64<0x141, 0xF>(v, 0.0);   // row_half_mirror
  v += dpp_f64<0x140, 0xF>(v, 0.0);   // row_mirror
  v += dpp_f64<0x142, 0xA>(v, 0.0);   // row_bcast:15 -> rows 1, 3
  v += dpp_f64<0x143, 0xC>(v, 0.0);   // row_bcast:31 -> rows 2, 3
  return lane63(v);
}
// Maximum over the wavefront, returned to every lane.
__device__ __forceinline__ double wave_max(double v) {
  const double lo = -__builtin_huge_val();
  double w;
  w = dpp_f64<0xB1, 0xF>(v, lo);  v = v > w ? v : w;
  w = dpp_f64<0x4E, 0xF>(v, lo);  v = v > w ? v : w;
  w = dpp_f64<0x141, 0xF>(v, lo); v = v > w ? v : w;
  w = dpp_f64<0x140, 0xF>(v, lo); v = v > w ? v : w;
  w = dpp_f64<0x142, 0xA>(v, lo); v = v > w ? v : w;
  w = dpp_f64<0x143, 0xC>(v, lo); v = v > w ? v : w;
  return lane63(v);
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
    constexpr int nw = TOPT_BLOCK / TOPT_WAVE;   // every kernel runs TOPT_BLOCK threads
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = sh[k];
      for (int w = 1; w < nw; ++w) t += sh[w * NV + k];
      v[k] = t;
    }
  }
  __syncthreads();
}

// Cross-block hand-off without an L2 flush per block. A block publishes its
// partials with agent-scope atomic stores (written through the XCD-local
// L2), waits until they are performed (s_waitcnt vmcnt(0)), then counts its
// arrival with an agent-scope atomic. Only the block that arrives last
// executes an agent-scope acquire (one L2 invalidate per kernel) and then
// reads the partials with ordinary loads, which the compiler batches.
// __threadfence() in every block would instead write back and invalidate
// the whole L2 once per block. Data other than the partials is only read by
// later kernels (kernel boundaries order it).
__device__ __forceinline__ void publish(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Called by one thread after its publish() calls; true for the last arrival.
__device__ __forceinline__ bool arrive(unsigned* counter, unsigned nb) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return t == nb - 1;
}
// Every thread of the last block calls this once before reading partials.
__device__ __forceinline__ void acquire_partials() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
__device__ __forceinline__ double collect(const double* p) { return gp(p)[0]; }
// The last block re-arms the counter for the next launch.
__device__ __forceinline__ void rearm(unsigned* counter) {
  __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Deterministic grid reduction over a 64-ary tree of blocks.
// Level 0 holds one partial per block; blocks [64g, 64g+64) form group g
// with its own arrival counter (counters 256 B apart, so at most 64
// atomics ever meet on one address). The last arriver of a group folds the
// group's partials in lane order with a wave-level tree, publishes the
// result one level up and arrives there; the block that folds the last
// level returns true with the totals. The summation order depends only on
// the number of blocks, never on arrival order or on other problems in a
// batch. maxmask: bit k set -> component k is combined with max, else sum.
// part: tree_parts(nb) * NV doubles; ctr: tree_ctrs(nb) * TOPT_CTR_STRIDE
// uints, zero between launches (every group re-arms its counter).
#define TOPT_TREE 64
#define TOPT_CTR_STRIDE 64

__host__ __device__ inline long tree_parts(long nb) {
  long t = 0, n = nb;
  for (;;) {
    t += n;
    if (n <= 1) return t;
    n = (n + TOPT_TREE - 1) / TOPT_TREE;
  }
}
__host__ __device__ inline long tree_ctrs(long nb) {
  long t = 0, n = nb;
  while (n > 1) {
    n = (n + TOPT_TREE - 1) / TOPT_TREE;
    t += n;
  }
  return t;
}

template <int NV>
__device__ bool tree_reduce(double (&v)[NV], double* part, unsigned* ctr, int lb, int nb,
                            double (&tot)[NV], unsigned maxmask = 0u) {
  __shared__ double tsh[NV];
  __shared__ int s_go;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int nw = TOPT_BLOCK / TOPT_WAVE;
  long idx = lb, n = nb;
  double* lvl = part;
  unsigned* cl = ctr;
  for (;;) {
    if (n <= 1) {
      if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) tsh[k] = v[k];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NV; ++k) tot[k] = tsh[k];
      __syncthreads();
      return true;
    }
    const long g = idx / TOPT_TREE, first = g * TOPT_TREE;
    const int cnt = (int)((n - first) < TOPT_TREE ? (n - first) : TOPT_TREE);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) publish(&lvl[idx * NV + k], v[k]);
      s_go = arrive(&cl[g * TOPT_CTR_STRIDE], (unsigned)cnt);
    }
    __syncthreads();
    if (!s_go) return false;
    acquire_partials();
    {
      // wave wv folds components wv, wv + nw, ...: all of its loads are
      // issued first (clamped addresses, the clamped values selected away),
      // then the folds
      constexpr int KW = (NV + nw - 1) / nw;
      const int ln = lane < cnt ? lane : cnt - 1;
      double xs[KW];
#pragma unroll
      for (int j = 0; j < KW; ++j) {
        const int k = wv + j * nw < NV ? wv + j * nw : NV - 1;
        xs[j] = collect(&lvl[(first + ln) * NV + k]);
      }
#pragma unroll
      for (int j = 0; j < KW; ++j) {
        const int k = wv + j * nw;
        if (k >= NV) break;
        const bool mx = (maxmask >> k) & 1u;
        double x = lane < cnt ? xs[j] : (mx ? -1.0e308 : 0.0);
        x = mx ? wave_max(x) : wave_sum(x);
        if (lane == 0) tsh[k] = x;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) v[k] = tsh[k];
      rearm(&cl[g * TOPT_CTR_STRIDE]);
    }
    __syncthreads();
    lvl += n * NV;
    cl += ((n + TOPT_TREE - 1) / TOPT_TREE) * TOPT_CTR_STRIDE;
    idx = g;
    n = (n + TOPT_TREE - 1) / TOPT_TREE;
  }
}

// Block sum followed by the grid tree (the usual pattern of the kernels).
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double* part, unsigned* ctr, int lb, int nb,
                                         double (&tot)[NV], double* sh) {
  block_sum<NV>(v, sh);
  return tree_reduce<NV>(v, part, ctr, lb, nb, tot);
}

// ------------------------------------------------------------ modal Q4 form
// On the unit-square bilinear element 2x2 Gauss integrates the strain energy
// exactly. For a stiffness symmetric under the square's reflections
// (isotropic plane stress is), the energy splits into the three constant
// strains and the two hourglass modes, which do not couple:
//   Ke p = G^T W G p,  G p = (A1, B2, A2 + B1, A3, B3),
// from differences of the corner values x_k, y_k (k = BL, BR, TR, TL):
//   A1 = (x_BR - x_BL) + (x_TR - x_TL)   twice the mean x-gradient of u_x
//   A2 = (x_TL - x_BL) + (x_TR - x_BR)   twice the mean y-gradient of u_x
//   A3 = (x_TR - x_TL) - (x_BR - x_BL)   hourglass amplitude of u_x
//   B1, B2, B3 the same for u_y,
// and W given by six weights km = (c11, c12, c22, c33, hx, hy):
//   q1 = c11 A1 + c12 B2, q2 = c12 A1 + c22 B2, q3 = c33 (A2 + B1),
//   tx = hx A3, ty = hy B3;
//   (Ke p)_BL = (tx - (q1 + q3),  ty - (q2 + q3))
//   (Ke p)_BR = ((q1 - q3) - tx,  -(q2 - q3) - ty)
//   (Ke p)_TR = ((q1 + q3) + tx,  (q2 + q3) + ty)
//   (Ke p)_TL = (-(q1 - q3) - tx, (q2 - q3) - ty).
// For plane stress with D = E/(1 - nu^2) [[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu)/2]]:
// c11 = c22 = D11/4, c12 = D12/4, c33 = D33/4, hx = hy = (D11 + D33)/12.
// About 40 operations per element instead of 64 multiply-adds, every one on
// differences of neighbouring values, so a common translation of the
// corners cancels in the first subtraction.
//
// The weights from four entries of Ke's first two rows (DOF order BLx, BLy,
// BRx, BRy, TRx, TRy, TLx, TLy): Ke[0][0] = c11 + c33 + hx,
// Ke[0][4] = hx - c11 - c33, Ke[0][1] = c12 + c33, Ke[0][3] = c12 - c33,
// Ke[1][1] = c22 + c33 + hy, Ke[1][5] = hy - c22 - c33.
inline void modal_weights(const double* Ke, double (&km)[6]) {
  const double c33 = 0.5 * (Ke[1] - Ke[3]);
  km[0] = 0.5 * (Ke[0] - Ke[4]) - c33;   // c11
  km[1] = 0.5 * (Ke[1] + Ke[3]);         // c12
  km[2] = 0.5 * (Ke[9] - Ke[13]) - c33;  // c22
  km[3] = c33;
  km[4] = 0.5 * (Ke[0] + Ke[4]);         // hx
  km[5] = 0.5 * (Ke[9] + Ke[13]);        // hy
}
// Ke p_e of one element in the modal form (host check of a given Ke).
inline void modal_apply_host(const double (&km)[6], const double* pe, double* out) {
  const double* x = pe;
  double xv[4], yv[4];
  for (int k = 0; k < 4; ++k) {
    xv[k] = x[2 * k];
    yv[k] = x[2 * k + 1];
  }
  const double A1 = (xv[1] - xv[0]) + (xv[2] - xv[3]), A2 = (xv[3] - xv[0]) + (xv[2] - xv[1]);
  const double A3 = (xv[2] - xv[3]) - (xv[1] - xv[0]);
  const double B1 = (yv[1] - yv[0]) + (yv[2] - yv[3]), B2 = (yv[3] - yv[0]) + (yv[2] - yv[1]);
  const double B3 = (yv[2] - yv[3]) - (yv[1] - yv[0]);
  const double q1 = km[0] * A1 + km[1] * B2, q2 = km[1] * A1 + km[2] * B2, q3 = km[3] * (A2 + B1);
  const double tx = km[4] * A3, ty = km[5] * B3;
  out[0] = tx - (q1 + q3);
  out[1] = ty - (q2 + q3);
  out[2] = (q1 - q3) - tx;
  out[3] = -(q2 - q3) - ty;
  out[4] = (q1 + q3) + tx;
  out[5] = (q2 + q3) + ty;
  out[6] = -(q1 - q3) - tx;
  out[7] = (q2 - q3) - ty;
}

// Cross-lane moves of the modal matvec (DPP wave shifts, VALU only): the
// value of lane + 1 (lane 63 reads 0) and of lane - 1 (lane 0 reads 0).
#define TOPT_DPP_WAVE_SHL1 0x130
#define TOPT_DPP_WAVE_SHR1 0x138
template <int CTRL>
__device__ __forceinline__ float wave_shift(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double wave_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Modal matvec walk geometry: a wave owns 64 consecutive node rows and
// stores the 62 inner ones (bands overlap by two rows); a workgroup holds
// four bands and one run of L node columns.
#define TOPT_MV_RB 62
__host__ __device__ inline int mv_band_groups(int nely) {
  const int bands = (nely + 1 + TOPT_MV_RB - 1) / TOPT_MV_RB;
  return (bands + TOPT_BLOCK / TOPT_WAVE - 1) / (TOPT_BLOCK / TOPT_WAVE);
}
__host__ __device__ inline long mv_blocks(int nelx, int nely, int L) {
  return (long)((nelx + 1 + L - 1) / L) * mv_band_groups(nely);
}

// ------------------------------------------------------------ Q4 stencil
// Node-centric gather of K(s) q at node (ix, iy) from the 3x3 neighbourhood
// values q[dx][dy] (dx, dy in 0..2 <-> offsets -1..+1) and the scales of the
// four adjacent elements sc[a][b] = s(ix-1+a, iy-1+b) (0 outside the grid).
// Element (ex, ey) has local nodes BL(0,0) BR(1,0) TR(1,1) TL(0,1); our node
// is its local node (1-a, 1-b).
template <typename T, typename KP>
__device__ __forceinline__ void q4_gather(KP K, const T (&sc)[2][2],
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
        tx = fma(K[(2 * me) * 8 + 2 * k], v.x, tx);
        tx = fma(K[(2 * me) * 8 + 2 * k + 1], v.y, tx);
        ty = fma(K[(2 * me + 1) * 8 + 2 * k], v.x, ty);
        ty = fma(K[(2 * me + 1) * 8 + 2 * k + 1], v.y, ty);
      }
      ax = fma(sc[a][b], tx, ax);
      ay = fma(sc[a][b], ty, ay);
    }
  }
  ox = ax;
  oy = ay;
}

// The unit-square Q4 plane-stress stiffness of material.py:20-86 has six
// distinct magnitudes whatever E and nu: Ke[i][j] = sign(KS[i][j]) *
// kc[|KS[i][j]|] (kc[0] is the diagonal). The host checks a given Ke against
// this pattern (kstruct); the structured stencil then keeps the six values
// in scalar registers for the whole kernel instead of reloading 64.
constexpr int KS[8][8] = {{0, 1, -2, -3, -4, -1, 5, 3},  {1, 0, 3, 5, -1, -4, -3, -2},
                          {-2, 3, 0, -1, 5, -3, -4, 1},  {-3, 5, -1, 0, 3, -2, 1, -4},
                          {-4, -1, 5, 3, 0, 1, -2, -3},  {-1, -4, -3, -2, 1, 0, 3, 5},
                          {5, -3, -4, 1, -2, 3, 0, -1},  {3, -2, 1, -4, -3, 5, -1, 0}};

// q4_gather with Ke from the structured pattern (FP64).
__device__ __forceinline__ void q4_gather_s(const double (&k)[6], const double (&sc)[2][2],
                                            const d2v (&q)[3][3], double& ox, double& oy) {
  constexpr int CX[4] = {0, 1, 1, 0};
  constexpr int CY[4] = {0, 0, 1, 1};
  double ax = 0.0, ay = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int me = (a == 1) ? ((b == 1) ? 0 : 3) : ((b == 1) ? 1 : 2);
      double tx = 0.0, ty = 0.0;
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const d2v v = q[a + CX[kk]][b + CY[kk]];
        const int i0 = KS[2 * me][2 * kk], i1 = KS[2 * me][2 * kk + 1];
        const int j0 = KS[2 * me + 1][2 * kk], j1 = KS[2 * me + 1][2 * kk + 1];
        tx = fma(i0 < 0 ? -k[-i0] : k[i0], v.x, tx);
        tx = fma(i1 < 0 ? -k[-i1] : k[i1], v.y, tx);
        ty = fma(j0 < 0 ? -k[-j0] : k[j0], v.x, ty);
        ty = fma(j1 < 0 ? -k[-j1] : k[j1], v.y, ty);
      }
      ax = fma(sc[a][b], tx, ax);
      ay = fma(sc[a][b], ty, ay);
    }
  }
  ox = ax;
  oy = ay;
}

// Scales of the four elements around node (ix, iy); zero outside the grid.
template <typename T>
__device__ __forceinline__ void load_scales(const T* s_generic, int ix, int iy, int nelx,
                                            int nely, T (&sc)[2][2]) {
  const GLB T* s = gp(s_generic);
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
