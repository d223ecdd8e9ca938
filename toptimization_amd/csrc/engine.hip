// libtopt host engine: workspace layout, batched plans, the PCG slot loop
// with asynchronous convergence polling, and the C ABI of include/topt.h.
//
// The reference drives every kernel from Python with a host sync after each
// scalar-returning kernel (fem/solver.py:107-156, optimizer/oc.py:71-88).
// Here alpha/beta/convergence/bisection live on the device; the host only
// enqueues fixed chunks of slot kernels and polls a status snapshot that is
// copied back asynchronously, one chunk behind the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/topt.h"
#include "topt_dev.h"
#include "fem_kernels.h"
#include "mg_kernels.h"
#include "opt_kernels.h"
#include "abi_internal.h"

using namespace topt;

// Scratch of topt_oc_update_noalloc (OC state, reduction counters, partial
// sums): device memory carried by the code object, one copy per GPU, held
// until the process ends. Host code never uses this name as an address: the
// per-device address comes from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[TOPT_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(TOPT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_TRY()                                                                     \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess) return fail(TOPT_EHIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct Geom {
  int64_t nn, ne, words;
  int nb[TOPT_NGRID];
  int ntree;   // blocks of the node-level reduction tree (node grid or tiles)
};

Geom geom(int nelx, int nely, int dtype = TOPT_F64) {
  Geom g;
  g.nn = int64_t(nelx + 1) * (nely + 1);
  g.ne = int64_t(nelx) * nely;
  g.words = cdiv(2 * g.nn, 32);
  g.nb[G_NODE] = (int)cdiv(g.nn, TOPT_BLOCK);
  g.nb[G_ELEM] = (int)cdiv(g.ne, TOPT_BLOCK);
  g.nb[G_FILT] = (int)(cdiv(nelx, TOPT_FILT_TX) * cdiv(nely, TOPT_FILT_TY));
  g.nb[G_OC] = (int)cdiv(g.ne, int64_t(TOPT_BLOCK) * TOPT_OC_PER_THREAD);
  g.nb[G_PCG] = (int)(dtype == TOPT_F32 ? pcg_tiles<float>(nelx, nely) : pcg_tiles<double>(nelx, nely));
  // partials of the widest reduction: node blocks, PCG tiles, op-level matvec tiles
  g.ntree = (int)std::max<int64_t>(
      std::max<int64_t>(g.nb[G_NODE], std::max(pcg_tiles<float>(nelx, nely), pcg_tiles<double>(nelx, nely))),
      std::max(std::max(pcg_tiles<float, NptMv<float>::v>(nelx, nely), pcg_tiles<double, NptMv<double>::v>(nelx, nely)),
               mv_blocks(nelx, nely, TOPT_MV_LMIN)));
  return g;
}

// Filter stencil in the reference's table order (optimizer/filter.py:58-74):
// dx outer, dy inner over [-ceil(r), ceil(r)], kept when w = rmin - sqrt(d2) > 0,
// weight stored as float32 (filter.py:53,71).
std::vector<Tap> build_taps(double rmin) {
  std::vector<Tap> taps;
  const int rc = (int)std::ceil(rmin);
  for (int dx = -rc; dx <= rc; ++dx)
    for (int dy = -rc; dy <= rc; ++dy) {
      const double w = rmin - std::sqrt((double)(dx * dx + dy * dy));
      if (w > 0.0) taps.push_back(Tap{dx, dy, (double)(float)w});
    }
  return taps;
}

int tap_radius(const std::vector<Tap>& taps) {
  int R = 0;
  for (const Tap& t : taps) R = std::max(R, std::max(std::abs(t.dx), std::abs(t.dy)));
  return R;
}

// Counter words of the three reduction trees (node, element, OC grid).
int64_t ctr_words(const Geom& g) {
  return (tree_ctrs(g.ntree) + tree_ctrs(g.nb[G_ELEM]) + tree_ctrs(g.nb[G_OC]) + 1) * TOPT_CTR_STRIDE;
}

void set_counters(ProbDev& P, unsigned* base, const Geom& g) {
  P.ctr[G_NODE] = base;
  P.ctr[G_ELEM] = base + tree_ctrs(g.ntree) * TOPT_CTR_STRIDE;
  P.ctr[G_OC] = P.ctr[G_ELEM] + tree_ctrs(g.nb[G_ELEM]) * TOPT_CTR_STRIDE;
  P.ctr[G_FILT] = nullptr;
}

struct Layout {
  int64_t off[TOPT_BUF_COUNT];
  int64_t total;
};

// Multigrid hierarchy of a grid: level sizes and buffer offsets inside the
// TOPT_BUF_MG region (level 0 keeps x0 and r1 there; its b is the PCG r and
// its 1/diag the pre buffer).
struct MgLayout {
  int nlev = 0, ncd = 0, lt = 1;
  int nx[TOPT_MG_MAXLEV], ny[TOPT_MG_MAXLEV];
  int64_t K[TOPT_MG_MAXLEV], x[TOPT_MG_MAXLEV], b[TOPT_MG_MAXLEV], r[TOPT_MG_MAXLEV], d[TOPT_MG_MAXLEV];
  int64_t st[TOPT_MG_MAXLEV];
  int64_t inv = 0, r0b = 0, total = 0;   // r0b: level 0's second residual buffer
};

int mg_layout(int nelx, int nely, int64_t ts, MgLayout& M) {
  M.nlev = 1;
  M.nx[0] = nelx;
  M.ny[0] = nely;
  for (;;) {
    if (M.nlev >= TOPT_MG_MAXLEV) return fail(TOPT_EINVAL, "grid too large for the multigrid hierarchy");
    const int l = M.nlev++;
    M.nx[l] = (M.nx[l - 1] + 1) / 2;
    M.ny[l] = (M.ny[l - 1] + 1) / 2;
    if (2 * int64_t(M.nx[l] + 1) * (M.ny[l] + 1) <= TOPT_MG_DENSE) break;
    if (M.nx[l] == M.nx[l - 1] && M.ny[l] == M.ny[l - 1])
      return fail(TOPT_EINVAL, "multigrid coarsening makes no progress");
  }
  int64_t o = 0;
  auto take = [&](int64_t bytes) { const int64_t at = o; o += align256(bytes); return at; };
  const int L = M.nlev - 1;
  // first level of the single-workgroup tail: small enough for one node
  // per thread, and all tail levels' vectors within its LDS buffers
  auto nodes = [&](int l) { return int64_t(M.nx[l] + 1) * (M.ny[l] + 1); };
  auto tail_sum = [&](int from) {
    int64_t t = 0;
    for (int l = from; l <= L; ++l) t += nodes(l);
    return t;
  };
  M.lt = 1;
  while (M.lt < L && (nodes(M.lt) > TOPT_MG_TAIL || tail_sum(M.lt) > TOPT_MG_TAIL_SUM)) ++M.lt;
  if (tail_sum(M.lt) > TOPT_MG_TAIL_SUM) return fail(TOPT_EINVAL, "multigrid tail exceeds its LDS buffers");
  // level 0 borrows b and 1/diag from the PCG (r, Jacobi diagonal); x is
  // stored at the coarsest level only (the sweeps form the pre-smoothed
  // iterate from b), r at levels 1 .. L-1 (the corrected iterate); the
  // sweep levels 1 .. L-1 also keep their node records
  for (int l = 0; l < M.nlev; ++l) {
    const int64_t nn = nodes(l), ne = int64_t(M.nx[l]) * M.ny[l];
    M.K[l] = l > 0 ? take(ne * 36 * ts) : -1;
    M.x[l] = l == L ? take(nn * 2 * ts) : -1;
    M.r[l] = l > 0 && l < L ? take(nn * 2 * ts) : -1;
    M.b[l] = l > 0 ? take(nn * 2 * ts) : -1;
    M.d[l] = l > 0 ? take(nn * 2 * ts) : -1;
    M.st[l] = l > 0 && l < L ? take(nn * TOPT_MG_ST * ts) : -1;
  }
  M.ncd = 2 * (M.nx[L] + 1) * (M.ny[L] + 1);
  M.inv = take(int64_t(M.ncd) * M.ncd * 8);
  M.r0b = take(nodes(0) * 2 * ts);
  M.total = o;
  return TOPT_OK;
}

int make_layout(int nelx, int nely, int dtype, int precond, double rmin, Layout& L) {
  if (nelx < 1 || nely < 1) return fail(TOPT_EINVAL, "grid must have nelx, nely >= 1");
  if (dtype != TOPT_F32 && dtype != TOPT_F64) return fail(TOPT_EINVAL, "dtype must be TOPT_F32 or TOPT_F64");
  if (precond < TOPT_PC_NONE || precond > TOPT_PC_MG) return fail(TOPT_EINVAL, "unknown preconditioner");
  if (!(rmin > 0.0)) return fail(TOPT_EINVAL, "rmin must be > 0");
  const Geom g = geom(nelx, nely, dtype);
  const int64_t ts = dtype == TOPT_F32 ? 4 : 8;
  const int64_t ntaps = (int64_t)build_taps(rmin).size();
  int64_t sz[TOPT_BUF_COUNT] = {0};
  sz[TOPT_BUF_U] = g.nn * 16;
  sz[TOPT_BUF_F] = g.nn * 16;
  sz[TOPT_BUF_FIX] = g.words * 4;
  sz[TOPT_BUF_RHO] = sz[TOPT_BUF_RHOF] = sz[TOPT_BUF_DC] = sz[TOPT_BUF_DCF] = sz[TOPT_BUF_HS] = g.ne * 8;
  sz[TOPT_BUF_X] = 0;   // the PCG iterate is the FP64 solution u on both paths
  sz[TOPT_BUF_R] = sz[TOPT_BUF_Z] = sz[TOPT_BUF_P0] = sz[TOPT_BUF_P1] = sz[TOPT_BUF_AP] = g.nn * 2 * ts;
  sz[TOPT_BUF_PRE] = (precond == TOPT_PC_JACOBI || precond == TOPT_PC_BLOCK2 || precond == TOPT_PC_MG)
                         ? g.nn * 2 * ts : 0;
  sz[TOPT_BUF_PRE2] = precond == TOPT_PC_BLOCK2 ? g.nn * ts : 0;
  sz[TOPT_BUF_S] = g.ne * ts;
  sz[TOPT_BUF_S64] = dtype == TOPT_F32 ? g.ne * 8 : 0;     // plain path: s64 aliases s
  sz[TOPT_BUF_TAPS] = std::max<int64_t>(ntaps, 1) * (int64_t)sizeof(Tap);
  sz[TOPT_BUF_CG] = 2 * sizeof(CgState);
  sz[TOPT_BUF_OC] = sizeof(OcState);
  // PCG slot partials: two parities x nb x 2 values
  sz[TOPT_BUF_PART] = std::max<int64_t>(tree_parts(g.ntree) * 2, 4 * int64_t(g.nb[G_PCG])) * 8;
  sz[TOPT_BUF_OPART] = std::max<int64_t>(tree_parts(g.nb[G_ELEM]), tree_parts(g.nb[G_OC]) * 32) * 8;
  sz[TOPT_BUF_CTR] = ctr_words(g) * 4;
  if (precond == TOPT_PC_MG) {
    MgLayout M;
    const int rc = mg_layout(nelx, nely, ts, M);
    if (rc) return rc;
    sz[TOPT_BUF_MG] = M.total;
  }
  int64_t o = 0;
  for (int i = 0; i < TOPT_BUF_COUNT; ++i) {
    L.off[i] = o;
    o += align256(sz[i]);
  }
  if (sz[TOPT_BUF_X] == 0) L.off[TOPT_BUF_X] = L.off[TOPT_BUF_U];
  if (sz[TOPT_BUF_S64] == 0) L.off[TOPT_BUF_S64] = L.off[TOPT_BUF_S];
  L.total = o;
  return TOPT_OK;
}

// 8x8 inverse by Gauss-Jordan with partial pivoting (FP64).
bool invert8(const double* A, double* Ainv) {
  double M[8][16];
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 16; ++j) M[i][j] = j < 8 ? A[i * 8 + j] : (j - 8 == i ? 1.0 : 0.0);
  for (int c = 0; c < 8; ++c) {
    int piv = c;
    for (int i = c + 1; i < 8; ++i)
      if (std::fabs(M[i][c]) > std::fabs(M[piv][c])) piv = i;
    if (M[piv][c] == 0.0) return false;
    if (piv != c)
      for (int j = 0; j < 16; ++j) std::swap(M[c][j], M[piv][j]);
    const double d = M[c][c];
    for (int j = 0; j < 16; ++j) M[c][j] /= d;
    for (int i = 0; i < 8; ++i) {
      if (i == c) continue;
      const double f = M[i][c];
      if (f != 0.0)
        for (int j = 0; j < 16; ++j) M[i][j] -= f * M[c][j];
    }
  }
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) Ainv[i * 8 + j] = M[i][8 + j];
  return true;
}

// Material part of a descriptor: Ke (FP32 + FP64) and the regularised
// lambda_max(D^-1 Ke) of an element stiffness (D its diagonal): the largest
// eigenvalue of D^-1/2 Ke D^-1/2 by cyclic Jacobi rotations (8 x 8, to
// machine precision). It bounds lambda_max(D^-1 K) of every assembled level
// operator, whatever the element scales, masks and Galerkin agglomeration.
double jacobi_bound(const double* Ke) {
  double M[8][8];
  for (int i = 0; i < 8; ++i) {
    if (!(Ke[i * 9] > 0.0)) return -1.0;
  }
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) M[i][j] = Ke[i * 8 + j] / std::sqrt(Ke[i * 9] * Ke[j * 9]);
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < 8; ++i)
      for (int j = i + 1; j < 8; ++j) off += M[i][j] * M[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < 8; ++p)
      for (int q = p + 1; q < 8; ++q) {
        if (std::fabs(M[p][q]) < 1e-300) continue;
        const double th = 0.5 * std::atan2(2.0 * M[p][q], M[q][q] - M[p][p]);
        const double c = std::cos(th), sn = std::sin(th);
        for (int k = 0; k < 8; ++k) {   // columns p, q
          const double mkp = M[k][p], mkq = M[k][q];
          M[k][p] = c * mkp - sn * mkq;
          M[k][q] = sn * mkp + c * mkq;
        }
        for (int k = 0; k < 8; ++k) {   // rows p, q
          const double mpk = M[p][k], mqk = M[q][k];
          M[p][k] = c * mpk - sn * mqk;
          M[q][k] = sn * mpk + c * mqk;
        }
      }
  }
  double lam = M[0][0];
  for (int i = 1; i < 8; ++i) lam = std::max(lam, M[i][i]);
  return lam;
}

// inverse used by the element-block preconditioner,
// Ki = inv(Ke + a I), a = 0.1 mean|diag Ke| (fem/fields.py:139-143).
int fill_material(ProbDev& P, const double* Ke, double E_min, double penal) {
  P.Emin = E_min;
  P.penal = penal;
  double a = 0.0;
  for (int i = 0; i < 8; ++i) a += std::fabs(Ke[i * 9]);
  a = 0.1 * a / 8.0;
  P.alpha_reg = a;
  double Kr[64], Ki[64];
  for (int i = 0; i < 64; ++i) {
    P.Ke64[i] = Ke[i];
    P.Ke32[i] = (float)Ke[i];
    Kr[i] = Ke[i] + ((i % 9) == 0 ? a : 0.0);
  }
  if (!invert8(Kr, Ki)) return fail(TOPT_EINVAL, "Ke + alpha I is singular");
  // multigrid smoother damping: a fixed fraction of the guaranteed range
  const double mu = jacobi_bound(Ke);
  P.mg_omega = mu > 0.0 ? TOPT_MG_OMEGA_FRAC * 2.0 / mu : 0.5;
  for (int i = 0; i < 64; ++i) {
    P.Ki64[i] = Ki[i];
    P.Ki32[i] = (float)Ki[i];
  }
  // structured stencil when Ke follows the unit-square Q4 pattern (to rounding)
  const double kc[6] = {Ke[0], Ke[1], -Ke[2], -Ke[3], -Ke[4], Ke[6]};
  double kmax = 0.0, dev = 0.0;
  for (int i = 0; i < 64; ++i) kmax = std::max(kmax, std::fabs(Ke[i]));
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) {
      const int t = KS[i][j];
      const double v = t < 0 ? -kc[-t] : kc[t];
      dev = std::max(dev, std::fabs(v - Ke[i * 8 + j]));
    }
  P.kstruct = (kmax > 0.0 && dev <= 1e-14 * kmax) ? 1 : 0;
  for (int t = 0; t < 6; ++t) P.kc[t] = kc[t];
  // modal form (k_matvec_modal): the six weights from four entries of the
  // first two rows, then the whole matrix rebuilt from them and compared
  double km[6];
  modal_weights(Ke, km);
  double mdev = 0.0;
  for (int j = 0; j < 8; ++j) {
    double e[8] = {0, 0, 0, 0, 0, 0, 0, 0}, col[8];
    e[j] = 1.0;
    modal_apply_host(km, e, col);
    for (int i = 0; i < 8; ++i) mdev = std::max(mdev, std::fabs(col[i] - Ke[i * 8 + j]));
  }
  P.kmodal = (kmax > 0.0 && mdev <= 1e-14 * kmax) ? 1 : 0;
  for (int t = 0; t < 6; ++t) P.km[t] = km[t];
  return TOPT_OK;
}

int fill_prob(ProbDev& P, const topt_problem& p, int dtype) {
  std::memset(&P, 0, sizeof(P));
  Layout L;
  int rc = make_layout(p.nelx, p.nely, dtype, p.precond, p.rmin, L);
  if (rc) return rc;
  if (p.ws == nullptr) return fail(TOPT_EINVAL, "problem workspace is NULL");
  const Geom g = geom(p.nelx, p.nely, dtype);
  P.nelx = p.nelx;
  P.nely = p.nely;
  for (int k = 0; k < TOPT_NGRID; ++k) P.nb[k] = g.nb[k];
  P.precond = p.precond;
  P.max_it = p.max_cg_iter;
  const std::vector<Tap> taps = build_taps(p.rmin);
  P.ntaps = (int)taps.size();
  P.R = tap_radius(taps);
  rc = fill_material(P, p.Ke, p.E_min, p.penal);
  if (rc) return rc;
  P.tol2 = (p.cg_tol * std::sqrt(p.fnorm2)) * (p.cg_tol * std::sqrt(p.fnorm2));
  P.delta2 = p.delta * p.delta;
  P.volfrac = p.volfrac;
  P.oc_tol = p.oc_tol;
  P.oc_max_iter = p.oc_max_iter;
  P.move = p.move;
  P.fnorm2 = p.fnorm2;
  char* w = (char*)p.ws;
  auto at = [&](int b) { return (void*)(w + L.off[b]); };
  P.x = at(TOPT_BUF_X);
  P.r = at(TOPT_BUF_R);
  P.r1 = P.r;   // a plan of multigrid problems only points it at the second buffer
  P.z = at(TOPT_BUF_Z);
  P.p0 = at(TOPT_BUF_P0);
  P.p1 = at(TOPT_BUF_P1);
  P.Ap = at(TOPT_BUF_AP);
  P.pre = at(TOPT_BUF_PRE);
  P.pre2 = at(TOPT_BUF_PRE2);
  P.s = at(TOPT_BUF_S);
  P.s64 = (double*)at(TOPT_BUF_S64);
  P.u64 = (double*)at(TOPT_BUF_U);
  P.f = (const double*)at(TOPT_BUF_F);
  P.fix = (const uint32_t*)at(TOPT_BUF_FIX);
  P.rho = (double*)at(TOPT_BUF_RHO);
  P.rhof = (double*)at(TOPT_BUF_RHOF);
  P.dc = (double*)at(TOPT_BUF_DC);
  P.dcf = (double*)at(TOPT_BUF_DCF);
  P.dv = nullptr;
  P.Hs = (double*)at(TOPT_BUF_HS);
  P.taps = (const Tap*)at(TOPT_BUF_TAPS);
  P.cg = (CgState*)at(TOPT_BUF_CG);
  P.part = (double*)at(TOPT_BUF_PART);
  P.oc = (OcState*)at(TOPT_BUF_OC);
  P.opart = (double*)at(TOPT_BUF_OPART);
  set_counters(P, (unsigned*)at(TOPT_BUF_CTR), g);
  if (p.precond == TOPT_PC_MG) {
    MgLayout M;
    rc = mg_layout(p.nelx, p.nely, dtype == TOPT_F32 ? 4 : 8, M);
    if (rc) return rc;
    char* m = (char*)at(TOPT_BUF_MG);
    P.mg_nlev = M.nlev;
    P.mg_ncd = M.ncd;
    P.mg_lt = M.lt;
    P.mg_inv = (void*)(m + M.inv);
    P.r1 = (void*)(m + M.r0b);
    for (int l = 0; l < M.nlev; ++l) {
      MgLevel& V = P.mg[l];
      V.nx = M.nx[l];
      V.ny = M.ny[l];
      V.K = l > 0 ? (void*)(m + M.K[l]) : nullptr;
      V.x = M.x[l] >= 0 ? (void*)(m + M.x[l]) : nullptr;
      V.r = M.r[l] >= 0 ? (void*)(m + M.r[l]) : nullptr;
      V.b = l > 0 ? (void*)(m + M.b[l]) : P.r;
      V.dinv = l > 0 ? (void*)(m + M.d[l]) : P.pre;
      V.st = M.st[l] >= 0 ? (void*)(m + M.st[l]) : nullptr;
    }
  }
  return TOPT_OK;
}

size_t filter_shmem(int R) {
  return size_t(TOPT_FILT_TX + 2 * R) * size_t(TOPT_FILT_TY + 2 * R) * sizeof(double);
}

// out: pinned host memory mapped into the device (read by the host after the
// event recorded behind this kernel)
__global__ void k_gather_status(const ProbDev* tab, int n, int parity, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CgState* st = tab[i].cg + parity;
  out[2 * i] = st->status;
  out[2 * i + 1] = st->it;
  __threadfence_system();
}

// abort_w (may be null): the plan's one-launch OC abort word; when set, every
// problem reports oc_steps = -1, which the host turns into an error.
__global__ void k_gather_stats(const ProbDev* tab, const int* idx, int n, topt_iter_stats* out,
                               const unsigned* abort_w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = idx ? idx[t] : t;
  const CgState* st = tab[i].cg;
  const OcState* o = tab[i].oc;
  topt_iter_stats s;
  s.compliance = o->c;
  s.volume = o->vol;
  s.change = o->change;
  s.lambda = o->last_mid;
  s.cg_status = st->status;
  s.n_cg = (st->status == ST_BREAK || st->status == ST_MAXIT) ? tab[i].max_it
           : (st->status == ST_ZERO ? 0 : st->it);
  s.oc_steps = (abort_w != nullptr && *abort_w != 0u) ? -1 : o->steps;
  s.refines = st->refines;
  s.iter = o->iter;
  s.done = o->done;
  out[i] = s;
}

// Bench support: put every problem back into CG mode so slot kernels can be
// timed back to back on the current vectors; k_cg_finish ends it again.
__global__ void k_cg_force_run(const ProbDev* tab, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CgState* st = tab[i].cg;
  st->status = ST_RUN;
  st->op = OP_C;          // next slot A finalises a "C" and runs A_CG
  st->cflag = CF_CG;
  st->amode = OP_A_CG;
  st->pending = 0;
  st->upend = 1;          // ... including the deferred u += alpha p, with alpha = 0
  st->alpha = 0.0;        // (the bytes move, u keeps its values)
  st->it = 0;
  if (!(st->rz > 0.0)) st->rz = 1.0;
}

__global__ void k_cg_finish(const ProbDev* tab, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tab[i].cg[0].status = ST_CONV;
  tab[i].cg[1].status = ST_CONV;
}

// After a solve: make cg[0] hold the final state (the last slot that covered
// problem i wrote cg[parity[i]]).
__global__ void k_cg_settle(const ProbDev* tab, int n, const int* parity) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || parity[i] == 0) return;
  tab[i].cg[0] = tab[i].cg[1];
}

__global__ void k_oc_out(const OcState* o, double* out) {
  out[0] = o->change;
  out[1] = o->vol;
  out[2] = o->last_mid;
  out[3] = (double)o->steps;
}

// Zero n 32-bit words (used instead of hipMemsetAsync where a caller may
// capture the launches into a graph).
__global__ void k_zero_words(unsigned* p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0u;
}

// Per-device address of `scratch`, looked up once per device with that
// device current.
#define TOPT_MAX_DEVICES 64
std::mutex g_scratch_mu;
void* g_scratch_ptr[TOPT_MAX_DEVICES] = {};

int noalloc_scratch(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= TOPT_MAX_DEVICES) return fail(TOPT_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) return fail(TOPT_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TOPT_OK;
}

// Layout of the OC bisection's scratch inside one buffer: OcState, the OC
// tree's arrival counters (these two are zeroed per call), then the partial
// sums (31 bisection volumes per block at each tree level).
struct OcScratch {
  int64_t ctr_off, part_off, zero_words, total;
};

OcScratch oc_scratch(int nb_oc) {
  OcScratch o;
  o.ctr_off = align256(sizeof(OcState));
  o.part_off = align256(o.ctr_off + (tree_ctrs(nb_oc) + 1) * TOPT_CTR_STRIDE * 4);
  o.zero_words = o.part_off / 4;
  o.total = o.part_off + align256(tree_parts(nb_oc) * 32 * 8);
  return o;
}

// The OC update's launches (optimizer/oc.py:52-93): state reset, the chain
// pass (at least one bisection step), ten depth-5 subtree passes (with it,
// the reference's <= 50 steps; passes after the stop return at once),
// commit, results.
int launch_oc_update(ProbDev& P, double* out, hipStream_t st) {
  constexpr int D = 5;
  hipLaunchKernelGGL((k_oc_reset<false>), dim3(1), dim3(64), 0, st, nullptr, nullptr, 1, P);
  hipLaunchKernelGGL((k_oc_eval<D, false, true>), dim3(P.nb[G_OC]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P);
  for (int pass = 0; pass < (50 + D - 1) / D; ++pass)
    hipLaunchKernelGGL((k_oc_eval<D, false>), dim3(P.nb[G_OC]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P);
  hipLaunchKernelGGL((k_oc_apply<false>), dim3(P.nb[G_OC]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P);
  LAUNCH_TRY();
  hipLaunchKernelGGL(k_oc_out, dim3(1), dim3(1), 0, st, P.oc, out);
  LAUNCH_TRY();
  return TOPT_OK;
}

}  // namespace

int topt_abi::set_error(int code, const std::string& msg) { return fail(code, msg); }

// ============================================================== plan
#define TOPT_SUBSLOTS 4
struct topt_plan {
  int n = 0;
  int dtype = TOPT_F64;
  bool three_slot = false;   // element-block Jacobi or multigrid: slots A, B, C
  bool mg = false;           // some problem uses multigrid: V-cycle between slots B and C
  bool all_mg = false;       // every problem does: slot C fused with the level-0 up sweep
  // multigrid V-cycle maps, (problem, block) pairs: down / up sweeps at level
  // l (problems with l < their tail level), restrictions l -> l+1 (problems
  // with l < their tail level; grid of level l+1), and one block per
  // multigrid problem for the tail
  int mg_lt_max = 0;
  int2* mg_dn[TOPT_MG_MAXLEV] = {};
  int2* mg_rs[TOPT_MG_MAXLEV] = {};
  int n_dn[TOPT_MG_MAXLEV] = {}, n_rs[TOPT_MG_MAXLEV] = {};
  int2* mg_tail = nullptr;
  int n_tail = 0;
  // one-problem plans: the last two down sweeps above the tail (levels dr2_l
  // and dr2_l + 1) in one launch of n_dr2 workgroups (k_mg_down2); -1: off
  int dr2_l = -1, n_dr2 = 0;
  // setup maps: Galerkin l -> l+1 (level l+1 elements), diagonal of level l
  int2* mg_ga[TOPT_MG_MAXLEV] = {};
  int2* mg_dg[TOPT_MG_MAXLEV] = {};
  int n_ga[TOPT_MG_MAXLEV] = {}, n_dg[TOPT_MG_MAXLEV] = {};
  bool kstruct = true;    // every problem's Ke follows the unit-square Q4 pattern
  // one-launch OC (k_oc_coop): its abort word and the number of its blocks
  // the device holds at once; dbar stays null where it cannot run
  unsigned* dbar = nullptr;
  int occ_cap = 0;
  int maxR = 0;
  int max_cg = 0;
  std::vector<ProbDev> host;
  ProbDev* dtab = nullptr;
  int2* dmap[TOPT_NGRID] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // (problem, block)
  int nblk[TOPT_NGRID] = {0, 0, 0, 0, 0};
  int64_t q = 0;          // PCG slot sequence number of the next launch
  int64_t na = 0;         // slot-A launches so far in this solve (p buffer parity)
  // (status, iteration) per problem after a chunk: pinned host slots mapped
  // into the device address space (dstat: their device addresses), written
  // directly (no copy operation behind each chunk) by the last slot kernel
  // of a captured chunk, else by k_gather_status
  int* dstat[2] = {nullptr, nullptr};
  int* hstat[2] = {nullptr, nullptr};
  topt_iter_stats* dstats = nullptr;
  topt_iter_stats* hstats = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // HIP graph of one full chunk of PCG slots (TOPT_GRAPH_CHUNK iterations),
  // captured on a private stream the first time it is needed, replayed with
  // hipGraphLaunch on the caller's stream
  bool use_graph = false;
  hipStream_t cap = nullptr;
  // one graph per status slot: the chunk's slots, the last one writing the
  // status snapshot into dstat[slot]
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t mgraph[2] = {nullptr, nullptr};       // multigrid plans: one TOPT_MG_CHUNK-iteration chunk
  hipGraphExec_t mgexec[2] = {nullptr, nullptr};
  hipGraph_t mgraph2[2] = {nullptr, nullptr};      // ... and a 2-iteration chunk (near the expected stop)
  hipGraphExec_t mgexec2[2] = {nullptr, nullptr};
  // CG iterations of each problem's last lockstep solve (-1: none yet)
  std::vector<int> last_cg;
  // lockstep SIMP iteration (topt_plan_iteration), per filter mode: the
  // launches before the PCG (density filter, solve setup) and after it
  // (sensitivity, dc filter, OC, statistics gather)
  hipGraph_t igraph[2][2] = {};
  hipGraphExec_t iexec[2][2] = {};
  // phase timing of topt_plan_iteration (bench support)
  bool timing = false;
  hipEvent_t tev[4] = {nullptr, nullptr, nullptr, nullptr};
  double tsum[3] = {0.0, 0.0, 0.0};
  int64_t tcalls = 0;
  // PCG grid map of (problem, local tile) pairs: the full map, and two
  // compacted copies (alternating, each filled from its own pinned staging
  // buffer) that leave out problems whose solve has ended
  int2* dpfull = nullptr;       // = dmap[G_PCG]
  int2* dpmap[2] = {nullptr, nullptr};
  int2* hpmap[2] = {nullptr, nullptr};
  const int2* pcur = nullptr;   // map of the next launches
  int npcur = 0;                // its block count
  int pslot = 0;
  std::vector<int2> hmap[TOPT_NGRID];   // host copies of the full maps
  std::vector<char> live;       // problem still in the PCG map
  // subset launches (asynchronous driver): rotating map / index / stats
  // staging slots; a slot used at one chunk boundary is reused two
  // boundaries later, when the chunk enqueued after it has completed
  int2* sub_d[TOPT_SUBSLOTS][TOPT_NGRID] = {};
  int2* sub_h[TOPT_SUBSLOTS][TOPT_NGRID] = {};
  int* sub_di[TOPT_SUBSLOTS] = {};
  int* sub_hi[TOPT_SUBSLOTS] = {};
  topt_iter_stats* sub_hs[TOPT_SUBSLOTS] = {};
  int* hpar = nullptr;          // pinned: parity of each problem's final state
  int* dpar = nullptr;
};

namespace {

// A batched launch over grid kind g: its block count and the block->problem
// map of that same kind (the map holds exactly nb entries).
struct GridRef {
  int nb;
  const int2* map;
};
GridRef grid_of(const topt_plan* pl, int g) { return GridRef{pl->nblk[g], pl->dmap[g]}; }

// Multigrid per-solve setup (Galerkin levels, level diagonals, coarsest
// inverse) of one problem (asynchronous driver: the problems whose solve
// starts at a chunk boundary) ...
template <typename T>
void mg_setup(const ProbDev& P, hipStream_t s) {
  const int L = P.mg_nlev - 1;
  auto nodes = [&](int l) { return (int)cdiv(int64_t(P.mg[l].nx + 1) * (P.mg[l].ny + 1), TOPT_BLOCK); };
  auto elems = [&](int l) { return (int)cdiv(int64_t(P.mg[l].nx) * P.mg[l].ny, TOPT_BLOCK); };
  hipLaunchKernelGGL((k_mg_galerkin<T, true, false>), dim3(elems(1)), dim3(TOPT_BLOCK), 0, s, nullptr, nullptr, P, 0);
  for (int l = 1; l < L; ++l)
    hipLaunchKernelGGL((k_mg_level<T, false>), dim3(elems(l + 1) + nodes(l)), dim3(TOPT_BLOCK), 0, s, nullptr,
                       nullptr, nullptr, P, l, elems(l + 1));
  hipLaunchKernelGGL((k_mg_diag<T, false>), dim3(nodes(L)), dim3(TOPT_BLOCK), 0, s, nullptr, nullptr, P, L);
  hipLaunchKernelGGL((k_mg_coarse_setup<T, false>), dim3(1), dim3(TOPT_MG_SETUP_BLOCK), 0, s, nullptr, nullptr, P);
}

// ... and of every multigrid problem of the plan through the per-level maps
// (lockstep solves): the Galerkin product 0 -> 1, then one launch per level
// l >= 1 for the product l -> l+1 (problems with a level l+1) and level l's
// diagonal and records (problems with a level l), then the coarsest inverses.
template <typename T>
void mg_setup_all(topt_plan* pl, hipStream_t s) {
  const ProbDev dummy{};
  if (pl->n_ga[0] > 0)
    hipLaunchKernelGGL((k_mg_galerkin<T, true, true>), dim3(pl->n_ga[0]), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                       pl->mg_ga[0], dummy, 0);
  for (int l = 1; l < TOPT_MG_MAXLEV; ++l) {
    const int nga = l + 1 < TOPT_MG_MAXLEV ? pl->n_ga[l] : 0, ndg = pl->n_dg[l];
    if (nga + ndg == 0) continue;
    hipLaunchKernelGGL((k_mg_level<T, true>), dim3(nga + ndg), dim3(TOPT_BLOCK), 0, s, pl->dtab, pl->mg_ga[l],
                       pl->mg_dg[l], dummy, l, nga);
  }
  hipLaunchKernelGGL((k_mg_coarse_setup<T, true>), dim3(pl->n_tail), dim3(TOPT_MG_SETUP_BLOCK), 0, s, pl->dtab,
                     pl->mg_tail, dummy);
}

// The PCG tile map of a launch: null (the identity) for a one-problem plan
// whose map covers its only problem.
const int2* pcg_map(const topt_plan* pl) {
  return pl->n == 1 && pl->pcur == pl->dpfull ? nullptr : pl->pcur;
}

// The V-cycle's launches (DIRECT: a one-problem plan, whose kernels take
// its state slot as an argument).
template <typename T, bool DIRECT>
void mg_vcycle_k(topt_plan* pl, int par, bool fused, hipStream_t s, bool down0) {
  const ProbDev* tab = pl->dtab;
  const int nt = pl->npcur;        // live problems' tiles (the PCG map)
  const int2* tmap = pcg_map(pl);
  // a one-problem plan's level maps are identity maps: passed as null, and
  // its state slot goes straight to the kernels
  auto lmap = [&](const int2* m) { return pl->n == 1 ? (const int2*)nullptr : m; };
  const CgState* cgd = DIRECT ? pl->host[0].cg + par : nullptr;
  // down sweeps fused with the restrictions: b_{l+1} from b_l (b_0 = r); in
  // a fused plan level 0's runs inside slot B (k_slot_b_mg)
  if (down0) {
    if (pl->kstruct)
      hipLaunchKernelGGL((k_mg_down_restrict0<T, true, DIRECT>), dim3(pl->n_rs[0]), dim3(TOPT_BLOCK), 0, s, cgd,
                         tab, lmap(pl->mg_rs[0]), par);
    else
      hipLaunchKernelGGL((k_mg_down_restrict0<T, false, DIRECT>), dim3(pl->n_rs[0]), dim3(TOPT_BLOCK), 0, s, cgd,
                         tab, lmap(pl->mg_rs[0]), par);
  }
  for (int l = 1; l < pl->mg_lt_max; ++l) {
    if (pl->n_rs[l] == 0) continue;
    if (l == pl->dr2_l) {   // levels l and l + 1 together
      hipLaunchKernelGGL((k_mg_down2<T, DIRECT>), dim3(pl->n_dr2), dim3(TOPT_BLOCK), 0, s, cgd, tab,
                         (const int2*)nullptr, l, par);
      ++l;
      continue;
    }
    hipLaunchKernelGGL((k_mg_down_restrict<T, DIRECT>), dim3(pl->n_rs[l]), dim3(TOPT_BLOCK), 0, s, cgd, tab,
                       lmap(pl->mg_rs[l]), l, par);
  }
  hipLaunchKernelGGL((k_mg_tail<T, DIRECT>), dim3(pl->n_tail), dim3(TOPT_MG_TAIL_BLOCK), 0, s, cgd, tab,
                     lmap(pl->mg_tail), par);
  for (int l = pl->mg_lt_max - 1; l >= 1; --l)
    if (pl->n_dn[l] > 0)
      hipLaunchKernelGGL((k_mg_up<T, DIRECT>), dim3(pl->n_dn[l]), dim3(TOPT_BLOCK), 0, s, cgd, tab, lmap(pl->mg_dn[l]),
                         l, par);
  if (fused) return;
  if (pl->kstruct)
    hipLaunchKernelGGL((k_mg_up0<T, true, DIRECT>), dim3(nt), dim3(TOPT_BLOCK), 0, s, cgd, tab, tmap, par);
  else
    hipLaunchKernelGGL((k_mg_up0<T, false, DIRECT>), dim3(nt), dim3(TOPT_BLOCK), 0, s, cgd, tab, tmap, par);
}
// fused: the level-0 down sweep runs inside slot B (k_slot_b_mg), the
// level-0 up sweep inside slot C (k_slot_c_mg)
template <typename T>
void mg_vcycle(topt_plan* pl, int par, bool fused, hipStream_t s, bool down0) {
  if (pl->n == 1)
    mg_vcycle_k<T, true>(pl, par, fused, s, down0);
  else
    mg_vcycle_k<T, false>(pl, par, fused, s, down0);
}

template <typename T>
// stat: when set, the chunk's last slot kernel writes the status snapshot there
int launch_slots(topt_plan* pl, int count, hipStream_t s, int* stat = nullptr) {
  const ProbDev dummy{};
  const int nb = pl->npcur;
  const int2* map = pcg_map(pl);
  if (nb == 0) {   // every solve has ended: the sequence numbers still advance
    pl->q += (int64_t)(pl->three_slot ? 3 : 2) * count;
    pl->na += count;
    return TOPT_OK;
  }
  for (int k = 0; k < count; ++k) {
    int* st_last = k == count - 1 ? stat : nullptr;
    // slot A number na reads p buffer (na & 1) and writes the other one
    const int pa = (int)(pl->na & 1);
    pl->na++;
    const int qa = (int)(pl->q++ & 1);
    if (pl->kstruct)
      hipLaunchKernelGGL((k_slot_a<T, true, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy, qa,
                         pa);
    else
      hipLaunchKernelGGL((k_slot_a<T, true, false>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy, qa,
                         pa);
    if (pl->all_mg && !TOPT_MG_FUSE_B) {
      // slot C fused with the level-0 up sweep only
      hipLaunchKernelGGL((k_slot_b<T, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy,
                         (int)(pl->q++ & 1), pa ^ 1, (int*)nullptr);
      mg_vcycle<T>(pl, (int)(pl->q & 1), true, s, true);
      if (pl->kstruct)
        hipLaunchKernelGGL((k_slot_c_mg<T, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map,
                           (int)(pl->q++ & 1), pa ^ 1, st_last);
      else
        hipLaunchKernelGGL((k_slot_c_mg<T, false>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map,
                           (int)(pl->q++ & 1), pa ^ 1, st_last);
      continue;
    }
    if (pl->all_mg) {
      // slot B fused with the level-0 down sweep, slot C with the level-0 up sweep
      if (pl->kstruct)
        hipLaunchKernelGGL((k_slot_b_mg<T, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map,
                           (int)(pl->q++ & 1), pa ^ 1);
      else
        hipLaunchKernelGGL((k_slot_b_mg<T, false>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map,
                           (int)(pl->q++ & 1), pa ^ 1);
      mg_vcycle<T>(pl, (int)(pl->q & 1), true, s, false);
      if (pl->kstruct)
        hipLaunchKernelGGL((k_slot_c_mg<T, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map,
                           (int)(pl->q++ & 1), pa ^ 1, st_last);
      else
        hipLaunchKernelGGL((k_slot_c_mg<T, false>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map,
                           (int)(pl->q++ & 1), pa ^ 1, st_last);
      continue;
    }
    hipLaunchKernelGGL((k_slot_b<T, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy,
                       (int)(pl->q++ & 1), pa ^ 1, pl->three_slot ? (int*)nullptr : st_last);
    if (pl->mg) mg_vcycle<T>(pl, (int)(pl->q & 1), false, s, true);
    if (pl->three_slot)
      hipLaunchKernelGGL((k_slot_c<T, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy,
                         (int)(pl->q++ & 1), pa ^ 1, st_last);
  }
  LAUNCH_TRY();
  return TOPT_OK;
}

#define TOPT_GRAPH_CHUNK 32
#ifndef TOPT_MG_CHUNK
#define TOPT_MG_CHUNK 4   // CG iterations per chunk of a multigrid solve
#endif

// The slot launches of one full chunk (count iterations), captured into a
// graph. Every full chunk starts at an even sequence number and an even
// slot-A count (the ramp-up chunks 4 + 4 + 8 + 16 = 32 iterations leave both
// even; multigrid chunks are 2 iterations of 3 slots), so its kernel
// arguments are the same each time and one graph serves the whole solve.
// Capture what `body` enqueues on the plan's private stream into a graph
// (kernels only) and instantiate it, once.
int capture_graph(topt_plan* pl, const std::function<int(hipStream_t)>& body, hipGraph_t* graph,
                  hipGraphExec_t* exec) {
  if (*exec) return TOPT_OK;
  if (!pl->cap) HIP_TRY(hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking));
  HIP_TRY(hipStreamBeginCapture(pl->cap, hipStreamCaptureModeThreadLocal));
  const int rc = body(pl->cap);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(pl->cap, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(TOPT_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  *graph = g;
  HIP_TRY(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
  return TOPT_OK;
}

// A full chunk's slot launches, the last one also writing the status
// snapshot into dstat[slot] (the state every problem's slot writes after the
// chunk: a full chunk, even start and even length in slots, leaves it at
// parity 0; chunk graphs cover every problem of the plan).
int chunk_graph(topt_plan* pl, int count, int slot, hipGraph_t* graph, hipGraphExec_t* exec) {
  if (*exec) return TOPT_OK;
  const int64_t q0 = pl->q, na0 = pl->na;
  const int rc = capture_graph(pl, [&](hipStream_t cs) {
    return pl->dtype == TOPT_F32 ? launch_slots<float>(pl, count, cs, pl->dstat[slot])
                                 : launch_slots<double>(pl, count, cs, pl->dstat[slot]);
  }, graph, exec);
  pl->q = q0;     // capturing launched nothing
  pl->na = na0;
  return rc;
}

int enqueue_chunk(topt_plan* pl, int count, int slot, hipStream_t s) {
  int rc;
  const bool whole = (pl->q & 1) == 0 && (pl->na & 1) == 0 && pl->pcur == pl->dpfull;
  const bool user_graph = pl->use_graph && count == TOPT_GRAPH_CHUNK && whole;
  // a multigrid iteration is ~28 launches: replaying them cuts the host's share
  const bool mg_graph = pl->mg && (count == TOPT_MG_CHUNK || count == 2) && whole;
  if (user_graph || mg_graph) {
    if (user_graph) {
      if ((rc = chunk_graph(pl, count, slot, &pl->graph[slot], &pl->gexec[slot]))) return rc;
      HIP_TRY(hipGraphLaunch(pl->gexec[slot], s));
    } else if (count == TOPT_MG_CHUNK) {
      if ((rc = chunk_graph(pl, count, slot, &pl->mgraph[slot], &pl->mgexec[slot]))) return rc;
      HIP_TRY(hipGraphLaunch(pl->mgexec[slot], s));
    } else {
      if ((rc = chunk_graph(pl, count, slot, &pl->mgraph2[slot], &pl->mgexec2[slot]))) return rc;
      HIP_TRY(hipGraphLaunch(pl->mgexec2[slot], s));
    }
    const int per_it = pl->three_slot ? 3 : 2;
    pl->q += (int64_t)per_it * count;
    pl->na += count;
  } else {
    rc = pl->dtype == TOPT_F32 ? launch_slots<float>(pl, count, s) : launch_slots<double>(pl, count, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather_status, dim3(cdiv(pl->n, 256)), dim3(256), 0, s, pl->dtab, pl->n,
                       (int)(pl->q & 1), pl->dstat[slot]);
    LAUNCH_TRY();
  }
  HIP_TRY(hipEventRecord(pl->ev[slot], s));
  return TOPT_OK;
}

template <typename T>
int launch_solve_setup(topt_plan* pl, int warm, int use_filtered, hipStream_t s) {
  const ProbDev dummy{};
  hipLaunchKernelGGL((k_scale<T, true>), dim3(grid_of(pl, G_ELEM).nb), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                     grid_of(pl, G_ELEM).map, dummy, use_filtered);
  hipLaunchKernelGGL((k_precond<T, true>), dim3(grid_of(pl, G_NODE).nb), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                     grid_of(pl, G_NODE).map, dummy, 0);
  if (pl->mg) mg_setup_all<T>(pl, s);
  hipLaunchKernelGGL((k_cg_prepare<T, true>), dim3(grid_of(pl, G_NODE).nb), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                     grid_of(pl, G_NODE).map, dummy, warm, 0);
  LAUNCH_TRY();
  return TOPT_OK;
}

int plan_filter(topt_plan* pl, int mode, hipStream_t s) {
  const ProbDev dummy{};
  hipLaunchKernelGGL((k_filter<true>), dim3(grid_of(pl, G_FILT).nb), dim3(TOPT_BLOCK), filter_shmem(pl->maxR), s,
                     pl->dtab, grid_of(pl, G_FILT).map, dummy, mode);
  LAUNCH_TRY();
  return TOPT_OK;
}

// Wait for a chunk's event by polling it: a chunk of multigrid iterations
// lasts a few hundred microseconds and the host must enqueue the next one
// before the GPU runs dry, which a blocking wait's wake-up does not always
// allow. Falls back to the blocking wait after TOPT_SPIN_WAIT_US.
#ifndef TOPT_SPIN_WAIT_US
#define TOPT_SPIN_WAIT_US 20000
#endif
hipError_t wait_event(hipEvent_t ev) {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e != hipErrorNotReady) return e;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(TOPT_SPIN_WAIT_US))
      return hipEventSynchronize(ev);
  }
}

// One-launch OC over the `nvb` logical blocks of `map`. False when the plan
// cannot use it; the caller then enqueues the multi-launch sequence, which
// gives the same bits.
bool oc_coop_launch(topt_plan* pl, const int2* map, int nvb, hipStream_t s) {
  if (pl->dbar == nullptr || nvb <= 0) return false;
  const int G = std::min(nvb, pl->occ_cap);
  if (cdiv(nvb, G) > TOPT_OCC_MAXK) return false;
  hipLaunchKernelGGL((k_oc_coop<5>), dim3(G), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, nvb, pl->dbar);
  return true;
}

int plan_oc(topt_plan* pl, hipStream_t s) {
  const ProbDev dummy{};
  constexpr int D = 5;
  if (oc_coop_launch(pl, grid_of(pl, G_OC).map, grid_of(pl, G_OC).nb, s)) {
    LAUNCH_TRY();
    return TOPT_OK;
  }
  hipLaunchKernelGGL((k_oc_reset<true>), dim3(cdiv(pl->n, 256)), dim3(256), 0, s, pl->dtab, nullptr, pl->n,
                     dummy);
  hipLaunchKernelGGL((k_oc_eval<D, true, true>), dim3(grid_of(pl, G_OC).nb), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                     grid_of(pl, G_OC).map, dummy);
  for (int pass = 0; pass < (50 + D - 1) / D; ++pass)
    hipLaunchKernelGGL((k_oc_eval<D, true>), dim3(grid_of(pl, G_OC).nb), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                       grid_of(pl, G_OC).map, dummy);
  hipLaunchKernelGGL((k_oc_apply<true>), dim3(grid_of(pl, G_OC).nb), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                     grid_of(pl, G_OC).map, dummy);
  LAUNCH_TRY();
  return TOPT_OK;
}

int plan_sens(topt_plan* pl, int use_filtered, hipStream_t s) {
  const ProbDev dummy{};
  hipLaunchKernelGGL((k_sens<true>), dim3(grid_of(pl, G_ELEM).nb), dim3(TOPT_BLOCK), 0, s, pl->dtab,
                     grid_of(pl, G_ELEM).map, dummy, use_filtered);
  LAUNCH_TRY();
  return TOPT_OK;
}

// The PCG chunks of a solve whose setup is enqueued. Returns once the status
// snapshots show every solve ended; the problems' final states are then
// settled into cg[0] in stream order (nothing is waited for after the last
// poll: the chunk still in flight and the settle run while the host goes on).
int plan_solve_chunks(topt_plan* pl, hipStream_t s) {
  int rc;
  pl->q = 0;
  pl->na = 0;
  pl->pcur = pl->dpfull;
  pl->npcur = pl->nblk[G_PCG];
  std::fill(pl->live.begin(), pl->live.end(), 1);
  // slots issued so far; a solve needs at most max_cg CG slots plus refresh slots
  const int64_t slot_cap = 4 * int64_t(pl->max_cg) + 4096;
  int64_t issued = 0;
  // multigrid solves take tens of iterations of ~100 us: short chunks keep
  // the overshoot past convergence small
  int chunk = pl->mg ? TOPT_MG_CHUNK : 4;
  const int chunk_max = pl->mg ? TOPT_MG_CHUNK : TOPT_GRAPH_CHUNK;
  // Multigrid: the host sees a stop one chunk late, and the launches of the
  // chunk in flight then run empty (~20 us per iteration). Once the chunk in
  // flight is expected to reach the CG count of the problems' previous
  // solves, chunks of 2 iterations halve that overshoot.
  int expect = -1;
  if (pl->mg)
    for (int i = 0; i < pl->n; ++i)
      if (pl->last_cg[i] >= 0) expect = expect < 0 ? pl->last_cg[i] : std::min(expect, pl->last_cg[i]);
  std::vector<int> seen(pl->n, 0);   // CG iterations of each problem as of the snapshot read last
  int inflight = 0;                  // CG iterations enqueued after the chunk whose snapshot was read
  for (int k = 0; k < 2; ++k) {
    rc = enqueue_chunk(pl, chunk, k, s);
    if (rc) return rc;
    issued += chunk;
  }
  inflight = chunk;
  int wait = 0;
  for (;;) {
    HIP_TRY(wait_event(pl->ev[wait]));
    // the snapshot is one chunk old: a problem that had stopped then was
    // still covered by the chunk in flight, so its final state is in
    // cg[q & 1] for the current q; from the next chunk on it is left out
    bool running = false, dropped = false;
    int it_min = -1;   // least CG count among the problems still running
    for (int i = 0; i < pl->n; ++i) {
      if (!pl->live[i]) continue;
      const int it_i = pl->hstat[wait][2 * i + 1];
      if (pl->hstat[wait][2 * i] == ST_RUN) {
        running = true;
        it_min = it_min < 0 ? it_i : std::min(it_min, it_i);
      } else if (!seen[i]) {
        pl->last_cg[i] = it_i;
      }
      if (pl->hstat[wait][2 * i] != ST_RUN) seen[i] = 1;
      if (pl->hstat[wait][2 * i] != ST_RUN && !pl->use_graph) {
        pl->live[i] = 0;
        pl->hpar[i] = (int)(pl->q & 1);
        dropped = true;
      }
    }
    if (!running) break;
    if (issued > slot_cap) return fail(TOPT_EINVAL, "PCG slot loop did not terminate");
    if (dropped) {
      const int k = pl->pslot ^= 1;
      int m = 0;
      for (const int2& e : pl->hmap[G_PCG])
        if (pl->live[e.x]) pl->hpmap[k][m++] = e;
      if (m > 0)
        HIP_TRY(hipMemcpyAsync(pl->dpmap[k], pl->hpmap[k], sizeof(int2) * m, hipMemcpyHostToDevice, s));
      pl->pcur = pl->dpmap[k];
      pl->npcur = m;
    }
    chunk = std::min(chunk * 2, chunk_max);
    if (pl->mg && expect >= 0 && it_min >= 0 && it_min + inflight + chunk > expect) chunk = 2;
    rc = enqueue_chunk(pl, chunk, wait, s);
    if (rc) return rc;
    issued += chunk;
    inflight = chunk;
    wait ^= 1;
  }
  // problems still in the map end with the chunk in flight, at q & 1 (hpar
  // is pinned: the copy reads it in stream order, and the host writes it
  // again only after polling a later chunk)
  for (int i = 0; i < pl->n; ++i)
    if (pl->live[i]) pl->hpar[i] = (int)(pl->q & 1);
  HIP_TRY(hipMemcpyAsync(pl->dpar, pl->hpar, sizeof(int) * pl->n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_cg_settle, dim3(cdiv(pl->n, 256)), dim3(256), 0, s, pl->dtab, pl->n, pl->dpar);
  LAUNCH_TRY();
  pl->q = 0;
  pl->na = 0;
  return TOPT_OK;
}

int plan_solve(topt_plan* pl, int warm, int use_filtered, int* n_iter, int* status, hipStream_t s) {
  int rc = pl->dtype == TOPT_F32 ? launch_solve_setup<float>(pl, warm, use_filtered, s)
                                 : launch_solve_setup<double>(pl, warm, use_filtered, s);
  if (rc) return rc;
  if ((rc = plan_solve_chunks(pl, s))) return rc;
  hipLaunchKernelGGL(k_gather_stats, dim3(cdiv(pl->n, 256)), dim3(256), 0, s, pl->dtab, nullptr, pl->n, pl->dstats, pl->dbar);
  LAUNCH_TRY();
  HIP_TRY(hipMemcpyAsync(pl->hstats, pl->dstats, sizeof(topt_iter_stats) * pl->n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int i = 0; i < pl->n; ++i) {
    if (n_iter) n_iter[i] = pl->hstats[i].n_cg;
    if (status) status[i] = pl->hstats[i].cg_status;
  }
  return TOPT_OK;
}

int gather_stats(topt_plan* pl, topt_iter_stats* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_stats, dim3(cdiv(pl->n, 256)), dim3(256), 0, s, pl->dtab, nullptr, pl->n, pl->dstats, pl->dbar);
  LAUNCH_TRY();
  HIP_TRY(hipMemcpyAsync(pl->hstats, pl->dstats, sizeof(topt_iter_stats) * pl->n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::memcpy(out, pl->hstats, sizeof(topt_iter_stats) * pl->n);
  return TOPT_OK;
}

// ---------------------------------------------------------------- asynchronous driver
// Subset launches: the (problem, block) pairs of grid g for the problems in
// `ids`, staged in slot `slot` and copied to the device in stream order.
int subset_grid(topt_plan* pl, int slot, int g, const std::vector<int>& ids, hipStream_t s, GridRef& out) {
  int m = 0;
  for (int i : ids) {
    const ProbDev& P = pl->host[i];
    std::memcpy(pl->sub_h[slot][g] + m, pl->hmap[g].data() + P.b0[g], sizeof(int2) * P.nb[g]);
    m += P.nb[g];
  }
  if (m > 0) HIP_TRY(hipMemcpyAsync(pl->sub_d[slot][g], pl->sub_h[slot][g], sizeof(int2) * m, hipMemcpyHostToDevice, s));
  out = GridRef{m, pl->sub_d[slot][g]};
  return TOPT_OK;
}

int subset_index(topt_plan* pl, int slot, const std::vector<int>& ids, hipStream_t s) {
  std::memcpy(pl->sub_hi[slot], ids.data(), sizeof(int) * ids.size());
  HIP_TRY(hipMemcpyAsync(pl->sub_di[slot], pl->sub_hi[slot], sizeof(int) * ids.size(), hipMemcpyHostToDevice, s));
  return TOPT_OK;
}

// Start of an OC iteration for `ids`: density filter (density mode), element
// scales, preconditioner, PCG state (auto_stop: a design whose stopping rule
// has been met gets status ST_DONE instead of a solve).
template <typename T>
int subset_setup(topt_plan* pl, const std::vector<int>& ids, int slot, int sens_filter, hipStream_t s) {
  const ProbDev dummy{};
  GridRef gf, ge, gn;
  int rc;
  if ((rc = subset_grid(pl, slot, G_FILT, ids, s, gf)) || (rc = subset_grid(pl, slot, G_ELEM, ids, s, ge)) ||
      (rc = subset_grid(pl, slot, G_NODE, ids, s, gn)))
    return rc;
  if (!sens_filter)
    hipLaunchKernelGGL((k_filter<true>), dim3(gf.nb), dim3(TOPT_BLOCK), filter_shmem(pl->maxR), s, pl->dtab, gf.map,
                       dummy, (int)F_DENSITY);
  hipLaunchKernelGGL((k_scale<T, true>), dim3(ge.nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, ge.map, dummy,
                     sens_filter ? 0 : 1);
  hipLaunchKernelGGL((k_precond<T, true>), dim3(gn.nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, gn.map, dummy, 0);
  for (int i : ids)
    if (pl->host[i].precond == PC_MG) mg_setup<T>(pl->host[i], s);
  hipLaunchKernelGGL((k_cg_prepare<T, true>), dim3(gn.nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, gn.map, dummy, 2, 1);
  LAUNCH_TRY();
  return TOPT_OK;
}

// End of an OC iteration for `ids` (their solves have ended and cg[0] holds
// the final state): sensitivity, dc filter, OC bisection and commit, then
// the statistics, copied to the slot's pinned buffer.
int subset_post(topt_plan* pl, const std::vector<int>& ids, int slot, int sens_filter, hipStream_t s) {
  const ProbDev dummy{};
  constexpr int D = 5;
  GridRef ge, gf, go;
  int rc;
  if ((rc = subset_grid(pl, slot, G_ELEM, ids, s, ge)) || (rc = subset_grid(pl, slot, G_FILT, ids, s, gf)) ||
      (rc = subset_grid(pl, slot, G_OC, ids, s, go)) || (rc = subset_index(pl, slot, ids, s)))
    return rc;
  const int n = (int)ids.size();
  hipLaunchKernelGGL((k_sens<true>), dim3(ge.nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, ge.map, dummy, sens_filter ? 0 : 1);
  hipLaunchKernelGGL((k_filter<true>), dim3(gf.nb), dim3(TOPT_BLOCK), filter_shmem(pl->maxR), s, pl->dtab, gf.map,
                     dummy, (int)(sens_filter ? F_SENS : F_DCAVG));
  if (!oc_coop_launch(pl, go.map, go.nb, s)) {
    hipLaunchKernelGGL((k_oc_reset<true>), dim3(cdiv(n, 256)), dim3(256), 0, s, pl->dtab, pl->sub_di[slot], n, dummy);
    hipLaunchKernelGGL((k_oc_eval<D, true, true>), dim3(go.nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, go.map, dummy);
    for (int pass = 0; pass < (50 + D - 1) / D; ++pass)
      hipLaunchKernelGGL((k_oc_eval<D, true>), dim3(go.nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, go.map, dummy);
    hipLaunchKernelGGL((k_oc_apply<true>), dim3(go.nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, go.map, dummy);
  }
  hipLaunchKernelGGL(k_gather_stats, dim3(cdiv(n, 256)), dim3(256), 0, s, pl->dtab, pl->sub_di[slot], n, pl->dstats, pl->dbar);
  LAUNCH_TRY();
  HIP_TRY(hipMemcpyAsync(pl->sub_hs[slot], pl->dstats, sizeof(topt_iter_stats) * pl->n, hipMemcpyDeviceToHost, s));
  return TOPT_OK;
}

// Rebuild the compacted PCG map from pl->live (alternating device buffers).
int rebuild_pcg_map(topt_plan* pl, hipStream_t s) {
  const int k = pl->pslot ^= 1;
  int m = 0;
  for (const int2& e : pl->hmap[G_PCG])
    if (pl->live[e.x]) pl->hpmap[k][m++] = e;
  if (m > 0) HIP_TRY(hipMemcpyAsync(pl->dpmap[k], pl->hpmap[k], sizeof(int2) * m, hipMemcpyHostToDevice, s));
  pl->pcur = pl->dpmap[k];
  pl->npcur = m;
  return TOPT_OK;
}

#define TOPT_ASYNC_CHUNK 16   // CG iterations per chunk (even: parities repeat)

// Every chunk is TOPT_ASYNC_CHUNK CG iterations, so q and na are even at each
// boundary: the state of every problem is in cg[0] there and a new solve
// starts on p buffer 0 whichever chunk it joins. The snapshot read after
// chunk c is one chunk old; a solve seen ended in it is still covered by
// chunk c + 1 (which applies its deferred u update), so its OC step and the
// next setup are enqueued after c + 1, and the new solve joins at c + 2.
int plan_advance(topt_plan* pl, int sens_filter, int max_chunks, int32_t* phase, topt_iter_stats* stats,
                 int32_t* done_now, int64_t* cg_now, hipStream_t s) {
  const int n = pl->n;
  for (int i = 0; i < n; ++i) {
    if (phase[i] < 0 || phase[i] > 3) return fail(TOPT_EINVAL, "phase must be 0..3");
    done_now[i] = 0;
    cg_now[i] = 0;
  }
  auto setup = [&](const std::vector<int>& ids, int slot) {
    return pl->dtype == TOPT_F32 ? subset_setup<float>(pl, ids, slot, sens_filter, s)
                                 : subset_setup<double>(pl, ids, slot, sens_filter, s);
  };
  struct Pending { std::vector<int> ids; int slot; int64_t ready; };
  std::vector<Pending> pend;
  std::vector<int64_t> join(n, 0);   // first chunk whose snapshot describes the current solve
  int slot = 0;
  auto take_slot = [&]() { const int k = slot; slot = (slot + 1) % TOPT_SUBSLOTS; return k; };
  bool became_done = false;
  auto read_stats = [&](const Pending& p) {
    for (int i : p.ids) {
      stats[i] = pl->sub_hs[p.slot][i];
      if (stats[i].oc_steps < 0) return fail(TOPT_EHIP, "OC bisection: grid barrier timed out");
      done_now[i] += 1;
      cg_now[i] += stats[i].n_cg;
      if (stats[i].done) {
        phase[i] = 3;
        pl->live[i] = 0;
        became_done = true;
      }
    }
    return (int)TOPT_OK;
  };
  pl->q = 0;
  pl->na = 0;
  int64_t cnext = 0;   // index of the next chunk to enqueue
  int rc;
  // ---- initial boundary: OC steps still pending, new OC iterations
  {
    std::vector<int> p2, s0;
    for (int i = 0; i < n; ++i) {
      if (phase[i] == 2) p2.push_back(i);
      if (phase[i] == 0 || phase[i] == 2) s0.push_back(i);
    }
    if (!p2.empty()) {
      const int k = take_slot();
      if ((rc = subset_post(pl, p2, k, sens_filter, s))) return rc;
      pend.push_back(Pending{p2, k, 0});
    }
    if (!s0.empty() && (rc = setup(s0, take_slot()))) return rc;
    for (int i : s0) phase[i] = 1;
    for (int i = 0; i < n; ++i) pl->live[i] = phase[i] == 1;
    if ((rc = rebuild_pcg_map(pl, s))) return rc;
  }
  int wait = 0;
  bool any_live = pl->npcur > 0;
  if (any_live) {
    for (int k = 0; k < 2; ++k) {
      if ((rc = enqueue_chunk(pl, pl->mg ? TOPT_MG_CHUNK : TOPT_ASYNC_CHUNK, (int)(cnext & 1), s))) return rc;
      ++cnext;
    }
  }
  int64_t chunks = 0;
  while (any_live) {
    const int64_t c = cnext - 2;   // chunk whose snapshot is read now
    HIP_TRY(wait_event(pl->ev[wait]));
    ++chunks;
    // statistics of OC steps that completed before chunk c
    for (auto it = pend.begin(); it != pend.end();) {
      if (it->ready <= c) {
        if ((rc = read_stats(*it))) return rc;
        it = pend.erase(it);
      } else {
        ++it;
      }
    }
    // solves that ended by chunk c: OC step and the next solve's setup now
    // (after chunk c + 1, which still covers them, in stream order)
    std::vector<int> fin;
    bool changed = became_done;
    for (int i = 0; i < n; ++i) {
      if (!pl->live[i] || join[i] > c) continue;
      const int st = pl->hstat[wait][2 * i];
      if (st == ST_RUN) continue;
      pl->live[i] = 0;
      changed = true;
      if (st == ST_DONE) {
        phase[i] = 3;
        became_done = true;
      } else {
        fin.push_back(i);
      }
    }
    const bool stop = became_done || chunks >= max_chunks;
    if (!fin.empty()) {
      const int k = take_slot();
      if ((rc = subset_post(pl, fin, k, sens_filter, s))) return rc;
      pend.push_back(Pending{fin, k, cnext});
      if (!stop) {
        if ((rc = setup(fin, take_slot()))) return rc;
        for (int i : fin) {
          pl->live[i] = 1;
          join[i] = cnext;
        }
      } else {
        for (int i : fin) phase[i] = 0;   // the OC step is enqueued; the next iteration starts later
      }
    }
    any_live = false;
    for (int i = 0; i < n; ++i) any_live = any_live || pl->live[i];
    if (stop || !any_live) break;
    if (changed && (rc = rebuild_pcg_map(pl, s))) return rc;
    if ((rc = enqueue_chunk(pl, pl->mg ? TOPT_MG_CHUNK : TOPT_ASYNC_CHUNK, wait, s))) return rc;
    ++cnext;
    wait ^= 1;
  }
  // Drain what is enqueued. A problem still live stays in phase 1 even if
  // its solve ended in the last chunk: no later slot A has applied its
  // deferred u update yet, so the next call's first chunk covers it again.
  HIP_TRY(hipStreamSynchronize(s));
  for (const Pending& p : pend)
    if ((rc = read_stats(p))) return rc;
  pl->q = 0;
  pl->na = 0;
  pl->pcur = pl->dpfull;
  pl->npcur = pl->nblk[G_PCG];
  return TOPT_OK;
}

// Single-problem descriptor for op-level calls: geometry, material and the
// reduction scratch carved from ws (partials + a CgState/OcState for counters).
struct OpScratch {
  int64_t part_off, opart_off, cg_off, oc_off, ctr_off, taps_off, total;
};

OpScratch op_scratch(int nelx, int nely) {
  const Geom g = geom(nelx, nely);
  OpScratch o;
  int64_t off = 0;
  o.part_off = off;
  off += align256(tree_parts(g.ntree) * 2 * 8);
  o.opart_off = off;
  off += align256(std::max<int64_t>(tree_parts(g.nb[G_ELEM]), tree_parts(g.nb[G_OC]) * 32) * 8);
  o.cg_off = off;
  off += align256(sizeof(CgState));
  o.oc_off = off;
  off += align256(sizeof(OcState));
  o.ctr_off = off;
  off += align256(ctr_words(g) * 4);
  o.taps_off = off;
  off += align256(int64_t(2 * 34 + 1) * (2 * 34 + 1) * (int64_t)sizeof(Tap));
  o.total = off;
  return o;
}

int op_prob(ProbDev& P, int nelx, int nely, void* ws, hipStream_t s) {
  std::memset(&P, 0, sizeof(P));
  if (nelx < 1 || nely < 1) return fail(TOPT_EINVAL, "grid must have nelx, nely >= 1");
  if (ws == nullptr) return fail(TOPT_EINVAL, "workspace is NULL");
  const Geom g = geom(nelx, nely);
  P.nelx = nelx;
  P.nely = nely;
  for (int k = 0; k < TOPT_NGRID; ++k) P.nb[k] = g.nb[k];
  const OpScratch o = op_scratch(nelx, nely);
  char* w = (char*)ws;
  P.part = (double*)(w + o.part_off);
  P.opart = (double*)(w + o.opart_off);
  P.cg = (CgState*)(w + o.cg_off);
  P.oc = (OcState*)(w + o.oc_off);
  set_counters(P, (unsigned*)(w + o.ctr_off), g);
  // reduction counters must start at zero
  HIP_TRY(hipMemsetAsync(P.cg, 0, sizeof(CgState), s));
  HIP_TRY(hipMemsetAsync(P.oc, 0, sizeof(OcState), s));
  HIP_TRY(hipMemsetAsync(P.ctr[G_NODE], 0, ctr_words(g) * 4, s));
  return TOPT_OK;
}

int grid_prob(ProbDev& P, const topt_grid* g, void* ws, hipStream_t s) {
  if (g == nullptr) return fail(TOPT_EINVAL, "grid is NULL");
  int rc = op_prob(P, g->nelx, g->nely, ws, s);
  if (rc) return rc;
  return fill_material(P, g->Ke, g->E_min, g->penal);
}

}  // namespace

// ============================================================== C ABI
extern "C" {

const char* topt_last_error(void) { return g_err.c_str(); }

int topt_device_count(int* n) {
  HIP_TRY(hipGetDeviceCount(n));
  return TOPT_OK;
}

int topt_problem_layout(int32_t nelx, int32_t nely, int32_t dtype, int32_t precond, double rmin,
                        int64_t* offsets, int64_t* total_bytes) {
  Layout L;
  const int rc = make_layout(nelx, nely, dtype, precond, rmin, L);
  if (rc) return rc;
  if (offsets)
    for (int i = 0; i < TOPT_BUF_COUNT; ++i) offsets[i] = L.off[i];
  if (total_bytes) *total_bytes = L.total;
  return TOPT_OK;
}

int topt_problem_init(const topt_problem* prob, int32_t dtype, void* stream) {
  if (prob == nullptr) return fail(TOPT_EINVAL, "problem is NULL");
  hipStream_t s = (hipStream_t)stream;
  ProbDev P;
  int rc = fill_prob(P, *prob, dtype);
  if (rc) return rc;
  const std::vector<Tap> taps = build_taps(prob->rmin);
  HIP_TRY(hipMemcpyAsync((void*)P.taps, taps.data(), taps.size() * sizeof(Tap), hipMemcpyHostToDevice, s));
  const Geom g = geom(prob->nelx, prob->nely);
  HIP_TRY(hipMemsetAsync(P.u64, 0, g.nn * 16, s));
  HIP_TRY(hipMemsetAsync(P.cg, 0, 2 * sizeof(CgState), s));
  HIP_TRY(hipMemsetAsync(P.oc, 0, sizeof(OcState), s));
  HIP_TRY(hipMemsetAsync(P.ctr[G_NODE], 0, ctr_words(g) * 4, s));
  P.b0[G_FILT] = 0;
  hipLaunchKernelGGL((k_filter<false>), dim3(g.nb[G_FILT]), dim3(TOPT_BLOCK), filter_shmem(P.R), s, nullptr,
                     nullptr, P, (int)F_HS);
  LAUNCH_TRY();
  HIP_TRY(hipStreamSynchronize(s));
  return TOPT_OK;
}

int topt_plan_create(const topt_problem* probs, int32_t nprob, int32_t dtype, topt_plan** out) {
  if (out == nullptr || probs == nullptr || nprob < 1) return fail(TOPT_EINVAL, "invalid plan arguments");
  topt_plan* pl = new topt_plan();
  pl->n = nprob;
  pl->dtype = dtype;
  pl->host.resize(nprob);
  std::vector<int> maps[TOPT_NGRID];
  for (int i = 0; i < nprob; ++i) {
    ProbDev& P = pl->host[i];
    const int rc = fill_prob(P, probs[i], dtype);
    if (rc) {
      delete pl;
      return rc;
    }
    for (int g = 0; g < TOPT_NGRID; ++g) {
      P.b0[g] = (int)maps[g].size();
      maps[g].insert(maps[g].end(), P.nb[g], i);
    }
    const bool three = P.precond == PC_ELEMBLOCK || P.precond == PC_MG;
    if (i > 0 && three != pl->three_slot) {
      delete pl;
      return fail(TOPT_EINVAL, "a plan cannot mix the element-block or multigrid preconditioner with "
                               "others (their PCG iteration has three slots)");
    }
    pl->three_slot = three;
    pl->mg = pl->mg || P.precond == PC_MG;
    pl->all_mg = (i == 0 || pl->all_mg) && P.precond == PC_MG;
    pl->kstruct = pl->kstruct && P.kstruct;
    pl->maxR = std::max(pl->maxR, P.R);
    pl->max_cg = std::max(pl->max_cg, P.max_it);
  }
  for (int g = 0; g < TOPT_NGRID; ++g) pl->nblk[g] = (int)maps[g].size();
  // the residual alternates between two buffers only where slot B is fused
  // with the level-0 down sweep
  if (!pl->all_mg || !TOPT_MG_FUSE_B)
    for (ProbDev& P : pl->host) P.r1 = P.r;
  auto cleanup = [&](int code) {
    topt_plan_destroy(pl);
    return code;
  };
  if (hipMalloc(&pl->dtab, sizeof(ProbDev) * nprob) != hipSuccess) return cleanup(fail(TOPT_ENOMEM, "hipMalloc plan table"));
  if (hipMemcpy(pl->dtab, pl->host.data(), sizeof(ProbDev) * nprob, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(TOPT_EHIP, "upload plan table"));
  for (int g = 0; g < TOPT_NGRID; ++g) {
    std::vector<int2>& m = pl->hmap[g];
    m.resize(maps[g].size());
    for (size_t b = 0; b < maps[g].size(); ++b) {
      const int i = maps[g][b];
      m[b] = make_int2(i, (int)b - pl->host[i].b0[g]);
    }
    if (hipMalloc(&pl->dmap[g], sizeof(int2) * m.size()) != hipSuccess)
      return cleanup(fail(TOPT_ENOMEM, "hipMalloc block map"));
    if (hipMemcpy(pl->dmap[g], m.data(), sizeof(int2) * m.size(), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(TOPT_EHIP, "upload block map"));
    for (int k = 0; k < TOPT_SUBSLOTS; ++k) {
      if (hipMalloc(&pl->sub_d[k][g], sizeof(int2) * m.size()) != hipSuccess)
        return cleanup(fail(TOPT_ENOMEM, "hipMalloc subset map"));
      if (hipHostMalloc(&pl->sub_h[k][g], sizeof(int2) * m.size(), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(TOPT_ENOMEM, "hipHostMalloc subset map"));
    }
  }
  for (int k = 0; k < TOPT_SUBSLOTS; ++k) {
    if (hipMalloc(&pl->sub_di[k], sizeof(int) * nprob) != hipSuccess)
      return cleanup(fail(TOPT_ENOMEM, "hipMalloc subset index"));
    if (hipHostMalloc(&pl->sub_hi[k], sizeof(int) * nprob, hipHostMallocDefault) != hipSuccess)
      return cleanup(fail(TOPT_ENOMEM, "hipHostMalloc subset index"));
    if (hipHostMalloc(&pl->sub_hs[k], sizeof(topt_iter_stats) * nprob, hipHostMallocDefault) != hipSuccess)
      return cleanup(fail(TOPT_ENOMEM, "hipHostMalloc subset stats"));
  }
  if (hipMalloc(&pl->dstats, sizeof(topt_iter_stats) * nprob) != hipSuccess) return cleanup(fail(TOPT_ENOMEM, "hipMalloc stats"));
  for (int k = 0; k < 2; ++k) {
    if (hipHostMalloc(&pl->hstat[k], sizeof(int) * 2 * nprob, hipHostMallocMapped) != hipSuccess)
      return cleanup(fail(TOPT_ENOMEM, "hipHostMalloc status"));
    if (hipHostGetDevicePointer((void**)&pl->dstat[k], pl->hstat[k], 0) != hipSuccess)
      return cleanup(fail(TOPT_EHIP, "hipHostGetDevicePointer status"));
    if (hipEventCreateWithFlags(&pl->ev[k], hipEventDisableTiming) != hipSuccess)
      return cleanup(fail(TOPT_EHIP, "hipEventCreate"));
  }
  if (hipHostMalloc(&pl->hstats, sizeof(topt_iter_stats) * nprob, hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(TOPT_ENOMEM, "hipHostMalloc stats"));
  // PCG grid: (problem, local tile) pairs in the order of the block map
  const int npcg = pl->nblk[G_PCG];
  pl->live.assign(nprob, 1);
  pl->last_cg.assign(nprob, -1);
  pl->dpfull = pl->dmap[G_PCG];
  for (int k = 0; k < 2; ++k) {
    if (hipMalloc(&pl->dpmap[k], sizeof(int2) * npcg) != hipSuccess)
      return cleanup(fail(TOPT_ENOMEM, "hipMalloc PCG map"));
    if (hipHostMalloc(&pl->hpmap[k], sizeof(int2) * npcg, hipHostMallocDefault) != hipSuccess)
      return cleanup(fail(TOPT_ENOMEM, "hipHostMalloc PCG map"));
  }
  if (hipHostMalloc(&pl->hpar, sizeof(int) * nprob, hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(TOPT_ENOMEM, "hipHostMalloc parity"));
  if (hipMalloc(&pl->dpar, sizeof(int) * nprob) != hipSuccess) return cleanup(fail(TOPT_ENOMEM, "hipMalloc parity"));
  pl->pcur = pl->dpfull;
  pl->npcur = npcg;
  if (pl->mg) {
    std::vector<int2> dn[TOPT_MG_MAXLEV], rs[TOPT_MG_MAXLEV], tail, ga[TOPT_MG_MAXLEV], dg[TOPT_MG_MAXLEV];
    for (int i = 0; i < nprob; ++i) {
      const ProbDev& P = pl->host[i];
      if (P.precond != PC_MG) continue;
      auto blocks = [&](int l) { return (int)cdiv(int64_t(P.mg[l].nx + 1) * (P.mg[l].ny + 1), TOPT_BLOCK); };
      // fused down sweeps + restrictions l -> l+1 for l < lt (tiles of
      // level l+1 nodes); up sweeps at levels 1 .. lt-1
      for (int l = 0; l < P.mg_lt; ++l) {
        if (l >= 1)
          for (int b = 0; b < mg_up_tiles(P.mg[l].nx, P.mg[l].ny); ++b) dn[l].push_back(make_int2(i, b));
        for (int b = 0; b < mg_dr_tiles(P.mg[l + 1].nx, P.mg[l + 1].ny); ++b) rs[l].push_back(make_int2(i, b));
      }
      tail.push_back(make_int2(i, 0));
      pl->mg_lt_max = std::max(pl->mg_lt_max, P.mg_lt);
      const int L = P.mg_nlev - 1;
      for (int l = 0; l < L; ++l) {
        const int ne = (int)cdiv(int64_t(P.mg[l + 1].nx) * P.mg[l + 1].ny, TOPT_BLOCK);
        for (int b = 0; b < ne; ++b) ga[l].push_back(make_int2(i, b));
      }
      for (int l = 1; l <= L; ++l)
        for (int b = 0; b < blocks(l); ++b) dg[l].push_back(make_int2(i, b));
    }
    auto upload = [&](const std::vector<int2>& v, int2** d, int* n) {
      *n = (int)v.size();
      if (v.empty()) return TOPT_OK;
      if (hipMalloc(d, sizeof(int2) * v.size()) != hipSuccess) return fail(TOPT_ENOMEM, "hipMalloc multigrid map");
      if (hipMemcpy(*d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(TOPT_EHIP, "upload multigrid map");
      return TOPT_OK;
    };
    int rc;
    for (int l = 0; l < TOPT_MG_MAXLEV; ++l)
      if ((rc = upload(dn[l], &pl->mg_dn[l], &pl->n_dn[l])) || (rc = upload(rs[l], &pl->mg_rs[l], &pl->n_rs[l])) ||
          (rc = upload(ga[l], &pl->mg_ga[l], &pl->n_ga[l])) || (rc = upload(dg[l], &pl->mg_dg[l], &pl->n_dg[l])))
        return cleanup(rc);
    if ((rc = upload(tail, &pl->mg_tail, &pl->n_tail))) return cleanup(rc);
    // the two-level down sweep: one-problem plans whose tail starts at level
    // 3 or deeper (TOPT_MG_NODOWN2 keeps the separate sweeps)
    const ProbDev& P0 = pl->host[0];
    if (nprob == 1 && P0.precond == PC_MG && P0.mg_lt >= 3 &&
        std::getenv("TOPT_MG_NODOWN2") == nullptr) {
      pl->dr2_l = P0.mg_lt - 2;
      pl->n_dr2 = mg_dr2_tiles(P0.mg[P0.mg_lt].nx, P0.mg[P0.mg_lt].ny);
    }
  }
  // one-launch OC: at least one of its blocks resident per CU
  {
    int dev = 0, ncu = 0, occ = 0;
    if (std::getenv("TOPT_OC_MULTI") == nullptr && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_oc_coop<5>, TOPT_BLOCK, 0) == hipSuccess &&
        occ >= 1 && ncu >= 1) {
      if (hipMalloc(&pl->dbar, TOPT_OCC_WORDS * 4) != hipSuccess) return cleanup(fail(TOPT_ENOMEM, "hipMalloc OC words"));
      if (hipMemset(pl->dbar, 0, TOPT_OCC_WORDS * 4) != hipSuccess) return cleanup(fail(TOPT_EHIP, "clear OC words"));
      // TOPT_OC_SPIN_TICKS: the limit of each wait in 100 MHz ticks instead
      // of 0.2 s (tests lower it to exercise the abort path)
      if (const char* e = std::getenv("TOPT_OC_SPIN_TICKS")) {
        const unsigned long long lim = std::strtoull(e, nullptr, 10);
        if (hipMemcpy(pl->dbar + 2, &lim, sizeof(lim), hipMemcpyHostToDevice) != hipSuccess)
          return cleanup(fail(TOPT_EHIP, "set OC spin limit"));
      }
      pl->occ_cap = std::min(TOPT_OCC_MAXG, occ * ncu);
    }
  }
  *out = pl;
  return TOPT_OK;
}

int topt_plan_destroy(topt_plan* pl) {
  if (pl == nullptr) return TOPT_OK;
  (void)hipFree(pl->dtab);
  (void)hipFree(pl->dbar);
  for (int g = 0; g < TOPT_NGRID; ++g) (void)hipFree(pl->dmap[g]);
  (void)hipFree(pl->dstats);
  for (int k = 0; k < 2; ++k) {
    if (pl->hstat[k]) (void)hipHostFree(pl->hstat[k]);
    if (pl->ev[k]) (void)hipEventDestroy(pl->ev[k]);
  }
  if (pl->hstats) (void)hipHostFree(pl->hstats);
  for (int k = 0; k < TOPT_SUBSLOTS; ++k) {
    for (int g = 0; g < TOPT_NGRID; ++g) {
      (void)hipFree(pl->sub_d[k][g]);
      if (pl->sub_h[k][g]) (void)hipHostFree(pl->sub_h[k][g]);
    }
    (void)hipFree(pl->sub_di[k]);
    if (pl->sub_hi[k]) (void)hipHostFree(pl->sub_hi[k]);
    if (pl->sub_hs[k]) (void)hipHostFree(pl->sub_hs[k]);
  }
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(pl->dpmap[k]);
    if (pl->hpmap[k]) (void)hipHostFree(pl->hpmap[k]);
  }
  if (pl->hpar) (void)hipHostFree(pl->hpar);
  (void)hipFree(pl->dpar);
  for (int k = 0; k < 2; ++k) {
    if (pl->gexec[k]) (void)hipGraphExecDestroy(pl->gexec[k]);
    if (pl->graph[k]) (void)hipGraphDestroy(pl->graph[k]);
    if (pl->mgexec[k]) (void)hipGraphExecDestroy(pl->mgexec[k]);
    if (pl->mgraph[k]) (void)hipGraphDestroy(pl->mgraph[k]);
    if (pl->mgexec2[k]) (void)hipGraphExecDestroy(pl->mgexec2[k]);
    if (pl->mgraph2[k]) (void)hipGraphDestroy(pl->mgraph2[k]);
    for (int j = 0; j < 2; ++j) {
      if (pl->iexec[k][j]) (void)hipGraphExecDestroy(pl->iexec[k][j]);
      if (pl->igraph[k][j]) (void)hipGraphDestroy(pl->igraph[k][j]);
    }
  }
  for (int k = 0; k < 4; ++k)
    if (pl->tev[k]) (void)hipEventDestroy(pl->tev[k]);
  for (int l = 0; l < TOPT_MG_MAXLEV; ++l) {
    (void)hipFree(pl->mg_dn[l]);
    (void)hipFree(pl->mg_rs[l]);
    (void)hipFree(pl->mg_ga[l]);
    (void)hipFree(pl->mg_dg[l]);
  }
  (void)hipFree(pl->mg_tail);
  if (pl->cap) (void)hipStreamDestroy(pl->cap);
  delete pl;
  return TOPT_OK;
}

int topt_plan_advance(topt_plan* pl, int32_t sens_filter, int32_t max_chunks, int32_t* phase,
                      topt_iter_stats* stats, int32_t* done_now, int64_t* cg_now, void* stream) {
  if (pl == nullptr || phase == nullptr || stats == nullptr || done_now == nullptr || cg_now == nullptr ||
      max_chunks < 1)
    return fail(TOPT_EINVAL, "invalid advance call");
  if (pl->use_graph) return fail(TOPT_EINVAL, "topt_plan_advance does not replay graphs");
  return plan_advance(pl, sens_filter, max_chunks, phase, stats, done_now, cg_now, (hipStream_t)stream);
}

int topt_plan_set_graph(topt_plan* pl, int32_t enable) {
  if (pl == nullptr) return fail(TOPT_EINVAL, "plan is NULL");
  pl->use_graph = enable != 0;
  return TOPT_OK;
}

int topt_plan_filter(topt_plan* pl, int32_t mode, void* stream) {
  if (pl == nullptr || mode < 0 || mode > 3) return fail(TOPT_EINVAL, "invalid filter call");
  return plan_filter(pl, mode, (hipStream_t)stream);
}

int topt_plan_solve(topt_plan* pl, int32_t warm, int32_t use_filtered, int32_t* n_iter, int32_t* status,
                    void* stream) {
  if (pl == nullptr) return fail(TOPT_EINVAL, "plan is NULL");
  return plan_solve(pl, warm, use_filtered, n_iter, status, (hipStream_t)stream);
}

int topt_plan_sensitivity(topt_plan* pl, int32_t use_filtered, void* stream) {
  if (pl == nullptr) return fail(TOPT_EINVAL, "plan is NULL");
  return plan_sens(pl, use_filtered, (hipStream_t)stream);
}

int topt_plan_oc(topt_plan* pl, void* stream) {
  if (pl == nullptr) return fail(TOPT_EINVAL, "plan is NULL");
  return plan_oc(pl, (hipStream_t)stream);
}

int topt_plan_iteration(topt_plan* pl, int32_t sens_filter, topt_iter_stats* stats, void* stream) {
  if (pl == nullptr) return fail(TOPT_EINVAL, "plan is NULL");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  const int m = sens_filter ? 1 : 0;
  const int use_f = sens_filter ? 0 : 1;
  // before the PCG: density filter and solve setup, one graph per mode
  rc = capture_graph(pl, [&](hipStream_t cs) {
    int r;
    if (!sens_filter && (r = plan_filter(pl, F_DENSITY, cs))) return r;
    return pl->dtype == TOPT_F32 ? launch_solve_setup<float>(pl, 2, use_f, cs)
                                 : launch_solve_setup<double>(pl, 2, use_f, cs);
  }, &pl->igraph[m][0], &pl->iexec[m][0]);
  if (rc) return rc;
  // after it: sensitivity, dc filter, OC bisection, statistics gather
  rc = capture_graph(pl, [&](hipStream_t cs) {
    int r;
    if ((r = plan_sens(pl, use_f, cs))) return r;
    if ((r = plan_filter(pl, sens_filter ? F_SENS : F_DCAVG, cs))) return r;
    if ((r = plan_oc(pl, cs))) return r;
    hipLaunchKernelGGL(k_gather_stats, dim3(cdiv(pl->n, 256)), dim3(256), 0, cs, pl->dtab, nullptr, pl->n,
                       pl->dstats, pl->dbar);
    LAUNCH_TRY();
    return (int)TOPT_OK;
  }, &pl->igraph[m][1], &pl->iexec[m][1]);
  if (rc) return rc;
  if (pl->timing) HIP_TRY(hipEventRecord(pl->tev[0], s));
  HIP_TRY(hipGraphLaunch(pl->iexec[m][0], s));
  if (pl->timing) HIP_TRY(hipEventRecord(pl->tev[1], s));
  if ((rc = plan_solve_chunks(pl, s))) return rc;
  if (pl->timing) HIP_TRY(hipEventRecord(pl->tev[2], s));
  HIP_TRY(hipGraphLaunch(pl->iexec[m][1], s));
  if (stats)
    HIP_TRY(hipMemcpyAsync(pl->hstats, pl->dstats, sizeof(topt_iter_stats) * pl->n, hipMemcpyDeviceToHost, s));
  if (pl->timing) HIP_TRY(hipEventRecord(pl->tev[3], s));
  if (!stats && !pl->timing) return TOPT_OK;
  HIP_TRY(hipStreamSynchronize(s));
  if (pl->timing) {
    for (int k = 0; k < 3; ++k) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, pl->tev[k], pl->tev[k + 1]));
      pl->tsum[k] += ms;
    }
    pl->tcalls++;
  }
  if (stats) {
    std::memcpy(stats, pl->hstats, sizeof(topt_iter_stats) * pl->n);
    for (int i = 0; i < pl->n; ++i)
      if (stats[i].oc_steps < 0) return fail(TOPT_EHIP, "OC bisection: grid barrier timed out");
  }
  return TOPT_OK;
}

int topt_plan_set_timing(topt_plan* pl, int32_t enable) {
  if (pl == nullptr) return fail(TOPT_EINVAL, "plan is NULL");
  for (int k = 0; k < 4 && enable; ++k)
    if (!pl->tev[k]) HIP_TRY(hipEventCreate(&pl->tev[k]));
  pl->timing = enable != 0;
  pl->tsum[0] = pl->tsum[1] = pl->tsum[2] = 0.0;
  pl->tcalls = 0;
  return TOPT_OK;
}

int topt_plan_timing(topt_plan* pl, double* ms3, int64_t* calls) {
  if (pl == nullptr || ms3 == nullptr) return fail(TOPT_EINVAL, "invalid timing call");
  for (int k = 0; k < 3; ++k) ms3[k] = pl->tsum[k];
  if (calls) *calls = pl->tcalls;
  return TOPT_OK;
}

int topt_plan_time_slot(topt_plan* pl, int32_t which, int32_t reps, float* ms, void* stream) {
  // Only slot A can be replayed: slot B would re-detect convergence and stop.
  if (pl == nullptr || ms == nullptr || reps < 1 || which != 0)
    return fail(TOPT_EINVAL, "invalid slot timing call (only which = 0 is supported)");
  hipStream_t s = (hipStream_t)stream;
  const ProbDev dummy{};
  const int nb = pl->nblk[G_PCG];
  const int2* map = pl->dpfull;
  // every launch re-reads cg[0] (forced to "after slot C") and repeats one A_CG step
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_cg_force_run, dim3(cdiv(pl->n, 256)), dim3(256), 0, s, pl->dtab, pl->n);
  HIP_TRY(hipEventRecord(e0, s));
  for (int k = 0; k < reps; ++k) {
    if (pl->dtype == TOPT_F32) {
      if (pl->kstruct)
        hipLaunchKernelGGL((k_slot_a<float, true, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy, 0, 0);
      else
        hipLaunchKernelGGL((k_slot_a<float, true, false>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy, 0, 0);
    } else {
      if (pl->kstruct)
        hipLaunchKernelGGL((k_slot_a<double, true, true>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy, 0, 0);
      else
        hipLaunchKernelGGL((k_slot_a<double, true, false>), dim3(nb), dim3(TOPT_BLOCK), 0, s, pl->dtab, map, dummy, 0, 0);
    }
  }
  HIP_TRY(hipEventRecord(e1, s));
  hipLaunchKernelGGL(k_cg_finish, dim3(cdiv(pl->n, 256)), dim3(256), 0, s, pl->dtab, pl->n);
  LAUNCH_TRY();
  HIP_TRY(hipEventSynchronize(e1));
  HIP_TRY(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return TOPT_OK;
}

int topt_plan_stats(topt_plan* pl, topt_iter_stats* stats, void* stream) {
  if (pl == nullptr || stats == nullptr) return fail(TOPT_EINVAL, "invalid stats call");
  return gather_stats(pl, stats, (hipStream_t)stream);
}

// ---------------------------------------------------------------- op level

int64_t topt_op_workspace_bytes(int32_t nelx, int32_t nely) {
  if (nelx < 1 || nely < 1) return -1;
  return op_scratch(nelx, nely).total;
}

int topt_scale(const topt_grid* g, int32_t dtype, const double* rho, void* s, void* stream) {
  if (g == nullptr || rho == nullptr || s == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  ProbDev P;
  std::memset(&P, 0, sizeof(P));
  P.nelx = g->nelx;
  P.nely = g->nely;
  P.Emin = g->E_min;
  P.penal = g->penal;
  P.rho = (double*)rho;
  P.s = s;
  const Geom gm = geom(g->nelx, g->nely);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TOPT_F32)
    hipLaunchKernelGGL((k_scale<float, false>), dim3(gm.nb[G_ELEM]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P, 0);
  else
    hipLaunchKernelGGL((k_scale<double, false>), dim3(gm.nb[G_ELEM]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P, 0);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt_matvec(const topt_grid* g, int32_t dtype, const void* p, void* Ap, const void* s, const uint32_t* fix,
                double* pAp, void* ws, void* stream) {
  if (p == nullptr || Ap == nullptr || s == nullptr || fix == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  ProbDev P;
  int rc;
  if (pAp != nullptr) {
    if ((rc = grid_prob(P, g, ws, st))) return rc;
  } else {
    if (g == nullptr) return fail(TOPT_EINVAL, "grid is NULL");
    std::memset(&P, 0, sizeof(P));
    P.nelx = g->nelx;
    P.nely = g->nely;
    const Geom gm = geom(g->nelx, g->nely);
    for (int k = 0; k < TOPT_NGRID; ++k) P.nb[k] = gm.nb[k];
    if ((rc = fill_material(P, g->Ke, g->E_min, g->penal))) return rc;
  }
  P.s = (void*)s;
  P.fix = fix;
  if (P.kmodal) {
    // modal walks of TOPT_MV_WALK node columns, or of the length the
    // environment variable TOPT_MV_WALK names (3, 7, 15, 31 or 63)
    int L = TOPT_MV_WALK;
    const char* fwalk = std::getenv("TOPT_MV_WALK");
    const int forced = fwalk ? std::atoi(fwalk) : 0;
    if (forced == 3 || forced == 7 || forced == 15 || forced == 31 || forced == 63) L = forced;
    P.nb[G_PCG] = (int)mv_blocks(P.nelx, P.nely, L);   // <= geom().ntree (L >= TOPT_MV_LMIN)
    const dim3 grid(P.nb[G_PCG]), blk(TOPT_BLOCK);
#define MV_MODAL(TT, LL)                                                                                  \
  do {                                                                                                    \
    if (pAp)                                                                                              \
      hipLaunchKernelGGL((k_matvec_modal<TT, LL, true>), grid, blk, 0, st, P, (const TT*)p, (TT*)Ap, pAp); \
    else                                                                                                  \
      hipLaunchKernelGGL((k_matvec_modal<TT, LL, false>), grid, blk, 0, st, P, (const TT*)p, (TT*)Ap, pAp); \
  } while (0)
#define MV_MODAL_L(TT)             \
  do {                             \
    if (L == 63) MV_MODAL(TT, 63); \
    else if (L == 31) MV_MODAL(TT, 31); \
    else if (L == 15) MV_MODAL(TT, 15); \
    else if (L == 7) MV_MODAL(TT, 7); \
    else MV_MODAL(TT, 3);          \
  } while (0)
    if (dtype == TOPT_F32) MV_MODAL_L(float);
    else MV_MODAL_L(double);
#undef MV_MODAL_L
#undef MV_MODAL
    LAUNCH_TRY();
    return TOPT_OK;
  }
  // one block per tile; the tree over those blocks (sized by geom().ntree)
  P.nb[G_PCG] = (int)(dtype == TOPT_F32 ? pcg_tiles<float, NptMv<float>::v>(P.nelx, P.nely)
                                         : pcg_tiles<double, NptMv<double>::v>(P.nelx, P.nely));
  const dim3 grid(P.nb[G_PCG]), blk(TOPT_BLOCK);
  if (dtype == TOPT_F32) {
    if (P.kstruct)
      hipLaunchKernelGGL((k_matvec<float, true>), grid, blk, 0, st, P, (const float*)p, (float*)Ap, pAp);
    else
      hipLaunchKernelGGL((k_matvec<float, false>), grid, blk, 0, st, P, (const float*)p, (float*)Ap, pAp);
  } else {
    if (P.kstruct)
      hipLaunchKernelGGL((k_matvec<double, true>), grid, blk, 0, st, P, (const double*)p, (double*)Ap, pAp);
    else
      hipLaunchKernelGGL((k_matvec<double, false>), grid, blk, 0, st, P, (const double*)p, (double*)Ap, pAp);
  }
  LAUNCH_TRY();
  return TOPT_OK;
}

static int precond_op(const topt_grid* g, int32_t dtype, const void* s, const uint32_t* fix, void* pre,
                      void* pre2, int kind, int raw, void* stream) {
  if (g == nullptr || s == nullptr || fix == nullptr || pre == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  ProbDev P;
  std::memset(&P, 0, sizeof(P));
  P.nelx = g->nelx;
  P.nely = g->nely;
  const Geom gm = geom(g->nelx, g->nely);
  for (int k = 0; k < TOPT_NGRID; ++k) P.nb[k] = gm.nb[k];
  int rc = fill_material(P, g->Ke, g->E_min, g->penal);
  if (rc) return rc;
  P.precond = kind;
  P.s = (void*)s;
  P.fix = fix;
  P.pre = pre;
  P.pre2 = pre2;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TOPT_F32)
    hipLaunchKernelGGL((k_precond<float, false>), dim3(gm.nb[G_NODE]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P, raw);
  else
    hipLaunchKernelGGL((k_precond<double, false>), dim3(gm.nb[G_NODE]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P, raw);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt_diagonal(const topt_grid* g, int32_t dtype, const void* s, const uint32_t* fix, void* diag, void* stream) {
  return precond_op(g, dtype, s, fix, diag, nullptr, PC_JACOBI, 1, stream);
}

int topt_block2_inverse(const topt_grid* g, int32_t dtype, const void* s, const uint32_t* fix, void* inv_diag,
                        void* inv_off, void* stream) {
  if (inv_off == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  return precond_op(g, dtype, s, fix, inv_diag, inv_off, PC_BLOCK2, 0, stream);
}

int topt_elemblock_apply(const topt_grid* g, int32_t dtype, const void* r, void* z, const void* s,
                         const uint32_t* fix, double* rz_rr, void* ws, void* stream) {
  if (r == nullptr || z == nullptr || s == nullptr || fix == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  ProbDev P;
  int rc = grid_prob(P, g, ws, st);
  if (rc) return rc;
  P.s = (void*)s;
  P.fix = fix;
  if (dtype == TOPT_F32)
    hipLaunchKernelGGL((k_elemblock<float>), dim3(P.nb[G_NODE]), dim3(TOPT_BLOCK), 0, st, P, (const float*)r,
                       (float*)z, rz_rr);
  else
    hipLaunchKernelGGL((k_elemblock<double>), dim3(P.nb[G_NODE]), dim3(TOPT_BLOCK), 0, st, P, (const double*)r,
                       (double*)z, rz_rr);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt_sensitivity(const topt_grid* g, const double* u, const double* rho, double* dc, double* dv,
                     double* compliance, void* ws, void* stream) {
  if (u == nullptr || rho == nullptr || dc == nullptr || compliance == nullptr)
    return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  ProbDev P;
  int rc = grid_prob(P, g, ws, st);
  if (rc) return rc;
  P.u64 = (double*)u;
  P.rho = (double*)rho;
  P.dc = dc;
  hipLaunchKernelGGL((k_sens<false>), dim3(P.nb[G_ELEM]), dim3(TOPT_BLOCK), 0, st, nullptr, nullptr, P, 0);
  LAUNCH_TRY();
  HIP_TRY(hipMemcpyAsync(compliance, &P.oc->c, sizeof(double), hipMemcpyDeviceToDevice, st));
  if (dv != nullptr) {
    std::vector<double> ones(size_t(g->nelx) * g->nely, 1.0);
    HIP_TRY(hipMemcpyAsync(dv, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return TOPT_OK;
}

int topt_filter(int32_t nelx, int32_t nely, double rmin, int32_t mode, const double* rho, const double* dc,
                double* out, double* Hs, void* ws, void* stream) {
  if (out == nullptr || Hs == nullptr || mode < 0 || mode > 3) return fail(TOPT_EINVAL, "invalid filter arguments");
  if (!(rmin > 0.0)) return fail(TOPT_EINVAL, "rmin must be > 0");
  hipStream_t st = (hipStream_t)stream;
  ProbDev P;
  int rc = op_prob(P, nelx, nely, ws, st);
  if (rc) return rc;
  const std::vector<Tap> taps = build_taps(rmin);
  if (tap_radius(taps) > 34) return fail(TOPT_EINVAL, "rmin too large for the op-level filter (max 34)");
  const OpScratch o = op_scratch(nelx, nely);
  P.taps = (const Tap*)((char*)ws + o.taps_off);
  P.ntaps = (int)taps.size();
  P.R = tap_radius(taps);
  HIP_TRY(hipMemcpyAsync((void*)P.taps, taps.data(), taps.size() * sizeof(Tap), hipMemcpyHostToDevice, st));
  P.rho = (double*)rho;
  P.dc = (double*)dc;
  P.Hs = Hs;
  P.rhof = out;
  P.dcf = out;
  hipLaunchKernelGGL((k_filter<false>), dim3(P.nb[G_FILT]), dim3(TOPT_BLOCK), filter_shmem(P.R), st, nullptr,
                     nullptr, P, (int)F_HS);
  hipLaunchKernelGGL((k_filter<false>), dim3(P.nb[G_FILT]), dim3(TOPT_BLOCK), filter_shmem(P.R), st, nullptr,
                     nullptr, P, (int)mode);
  LAUNCH_TRY();
  HIP_TRY(hipStreamSynchronize(st));   // taps upload came from pageable memory
  return TOPT_OK;
}

int topt_filter_tables(int32_t nelx, int32_t nely, double rmin, int32_t width, int32_t* nb, float* wt,
                       int32_t* cnt, void* ws, void* stream) {
  if (nb == nullptr || wt == nullptr || cnt == nullptr || ws == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  if (nelx < 1 || nely < 1 || !(rmin > 0.0)) return fail(TOPT_EINVAL, "invalid grid or rmin");
  const std::vector<Tap> taps = build_taps(rmin);
  if ((int)taps.size() > width) return fail(TOPT_EINVAL, "table width smaller than the stencil");
  if (tap_radius(taps) > 34) return fail(TOPT_EINVAL, "rmin too large (max 34)");
  hipStream_t st = (hipStream_t)stream;
  const OpScratch o = op_scratch(nelx, nely);
  Tap* dt = (Tap*)((char*)ws + o.taps_off);
  HIP_TRY(hipMemcpyAsync(dt, taps.data(), taps.size() * sizeof(Tap), hipMemcpyHostToDevice, st));
  const int64_t ne = int64_t(nelx) * nely;
  hipLaunchKernelGGL(k_filter_tables, dim3(cdiv(ne, TOPT_BLOCK)), dim3(TOPT_BLOCK), 0, st, nelx, nely, dt,
                     (int)taps.size(), width, nb, wt, cnt);
  LAUNCH_TRY();
  HIP_TRY(hipStreamSynchronize(st));
  return TOPT_OK;
}

int topt_oc_update(int32_t nelx, int32_t nely, double* rho, const double* dcf, const double* dv, double volfrac,
                   double move, double* out, void* ws, void* stream) {
  if (rho == nullptr || dcf == nullptr || out == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  ProbDev P;
  int rc = op_prob(P, nelx, nely, ws, st);
  if (rc) return rc;
  P.rho = rho;
  P.dcf = (double*)dcf;
  P.dv = dv;
  P.volfrac = volfrac;
  P.move = move;
  if (dv != nullptr) {
    // non-unit dv: exact per-lambda evaluation is not implemented in the
    // subtree pass; reject instead of silently assuming dv = 1
    return fail(TOPT_EINVAL, "topt_oc_update: dv must be NULL (dv = 1, fem/kernels.py:270)");
  }
  return launch_oc_update(P, out, st);
}

int topt_oc_update_noalloc(int32_t nelx, int32_t nely, double* rho, const double* dcf, const double* dv,
                           double volfrac, double move, double* out, void* stream) {
  if (rho == nullptr || dcf == nullptr || out == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  if (dv != nullptr) return fail(TOPT_EINVAL, "topt_oc_update_noalloc: dv must be NULL (dv = 1, fem/kernels.py:270)");
  if (nelx < 1 || nely < 1) return fail(TOPT_EINVAL, "grid must have nelx, nely >= 1");
  const Geom g = geom(nelx, nely);
  const OcScratch o = oc_scratch(g.nb[G_OC]);
  if (o.total > TOPT_NOALLOC_SCRATCH_BYTES)
    return fail(TOPT_EINVAL, "topt_oc_update_noalloc: the grid needs " + std::to_string(o.total) +
                                 " bytes of scratch, more than the static " +
                                 std::to_string((long long)TOPT_NOALLOC_SCRATCH_BYTES) +
                                 " (at most about 32 million elements); use topt_oc_update");
  void* base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  ProbDev P;
  std::memset(&P, 0, sizeof(P));
  P.nelx = nelx;
  P.nely = nely;
  for (int k = 0; k < TOPT_NGRID; ++k) P.nb[k] = g.nb[k];
  char* w = (char*)base;
  P.oc = (OcState*)w;
  P.ctr[G_OC] = (unsigned*)(w + o.ctr_off);
  P.opart = (double*)(w + o.part_off);
  P.rho = rho;
  P.dcf = (double*)dcf;
  P.dv = nullptr;
  P.volfrac = volfrac;
  P.move = move;
  // the state and the counters start at zero (a fill kernel, no memset node)
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<int64_t>(64, cdiv(o.zero_words, TOPT_BLOCK))),
                     dim3(TOPT_BLOCK), 0, st, (unsigned*)w, (long)o.zero_words);
  LAUNCH_TRY();
  return launch_oc_update(P, out, st);
}

int topt_volume(int32_t nelx, int32_t nely, const double* v, double* out, void* ws, void* stream) {
  if (v == nullptr || out == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  ProbDev P;
  int rc = op_prob(P, nelx, nely, ws, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_volume, dim3(P.nb[G_OC]), dim3(TOPT_BLOCK), 0, st, P, v, out);
  LAUNCH_TRY();
  return TOPT_OK;
}

// Vector ops of the drop-in API (fem/kernels.py:19-73,330-345,
// fem/fused_kernels.py:110-187, fem/solver.py:162-167). Buffers by role:
// x, y, z, w, v as documented per op in include/topt.h; red[2] (device)
// receives reductions. ws: >= topt_blas1_workspace_bytes().
int64_t topt_blas1_workspace_bytes(void) {
  return align256(tree_parts(1024) * 2 * 8) + tree_ctrs(1024) * TOPT_CTR_STRIDE * 4;
}

int topt_blas1(int32_t op, int32_t dtype, int64_t n, double a, void* x, void* y, void* z, void* w, void* v,
               const void* d, const uint32_t* fix, double* red, void* ws, void* stream) {
  if (op < B1_FILL || op > B1_ZERO_FIXED || n < 0) return fail(TOPT_EINVAL, "invalid blas1 op");
  if (n == 0) return TOPT_OK;
  const bool reduce = op == B1_DOT || op == B1_MAXDIFF || op == B1_SUM || op == B1_UPD_JAC || op == B1_UPD_NONE;
  if (reduce && (red == nullptr || ws == nullptr)) return fail(TOPT_EINVAL, "reduction needs red and ws");
  if (op == B1_ZERO_FIXED && fix == nullptr) return fail(TOPT_EINVAL, "zero_fixed needs the bitmask");
  if (op == B1_UPD_JAC && d == nullptr) return fail(TOPT_EINVAL, "jacobi update needs the diagonal");
  hipStream_t s = (hipStream_t)stream;
  const int grid = (int)std::min<int64_t>(1024, cdiv(n, TOPT_BLOCK));
  double* part = (double*)ws;
  unsigned* counter = (unsigned*)((char*)ws + align256(tree_parts(1024) * 2 * 8));
  if (reduce) HIP_TRY(hipMemsetAsync(counter, 0, tree_ctrs(1024) * TOPT_CTR_STRIDE * 4, s));
  if (dtype == TOPT_F32)
    hipLaunchKernelGGL((k_blas1<float>), dim3(grid), dim3(TOPT_BLOCK), 0, s, (int)op, (long)n, a, (float*)x,
                       (float*)y, (float*)z, (float*)w, (float*)v, (const float*)d, fix, part, counter, red);
  else
    hipLaunchKernelGGL((k_blas1<double>), dim3(grid), dim3(TOPT_BLOCK), 0, s, (int)op, (long)n, a, (double*)x,
                       (double*)y, (double*)z, (double*)w, (double*)v, (const double*)d, fix, part, counter, red);
  LAUNCH_TRY();
  return TOPT_OK;
}

}  // extern "C"
