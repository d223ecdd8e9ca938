// libtopt, 3-D H8 path: the C ABI of the topt3_* entry points (FP64).
//
// Op-level calls mirror the reference's 3-D Taichi kernels one to one; the
// solve runs the whole PCG of fem/solver.py:27-167 on the device (alpha,
// beta and the stopping tests in a device state, three or four launches per
// iteration, no host round trip per iteration): the host enqueues chunks of
// iterations and polls the state one chunk behind.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/topt.h"
#include "abi_internal.h"
#include "h8_kernels.h"

using namespace topt;

namespace {

int fail(int code, const std::string& msg) { return topt_abi::set_error(code, msg); }

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

constexpr int kMaxTapRadius = 12;   // (2*12+1)^3 = 15,625 candidate taps

__global__ void k3_zero_words(unsigned* p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k3_fill(double* p, long n, double v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// u[fixed] = 0 (solver.py:162-167)
__global__ void k3_zero_fixed(double* u, const uint32_t* fix, long nd) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nd; i += (long)gridDim.x * blockDim.x)
    if ((fix[i >> 5] >> (unsigned)(i & 31)) & 1u) u[i] = 0.0;
}

struct Dims3 {
  int64_t nn, ne, nd;
  int nbn, nbe;
  int nbt;   // tiles of the matvec (k3_matvec)
};

Dims3 dims3(int nelx, int nely, int nelz) {
  Dims3 d;
  d.nn = int64_t(nelx + 1) * (nely + 1) * (nelz + 1);
  d.ne = int64_t(nelx) * nely * nelz;
  d.nd = 3 * d.nn;
  d.nbn = (int)cdiv(d.nn, TOPT_BLOCK);
  d.nbe = (int)cdiv(d.ne, TOPT_BLOCK);
  d.nbt = mv3_tiles(nelx, nely, nelz);
  return d;
}

// Workspace: reduction partials and counters, PCG state, filter taps, PCG
// vectors r, z, p, Ap, diag.
struct Ws3 {
  int64_t part, ctr, cg, taps, r, z, p, Ap, diag, total, ctr_words;
};

Ws3 ws3(const Dims3& d) {
  Ws3 w;
  int64_t o = 0;
  auto take = [&](int64_t bytes) { const int64_t at = o; o += align256(bytes); return at; };
  // the node-grid and tile-grid reductions share the partials and counters
  const int64_t nbr = std::max<int64_t>(d.nbn, d.nbt);
  w.part = take(tree_parts(nbr) * 2 * 8);
  w.ctr_words = (tree_ctrs(nbr) + 1) * TOPT_CTR_STRIDE;
  w.ctr = take(w.ctr_words * 4);
  w.cg = take(sizeof(Cg3));
  w.taps = take(int64_t(2 * kMaxTapRadius + 1) * (2 * kMaxTapRadius + 1) * (2 * kMaxTapRadius + 1) * sizeof(Tap3));
  w.r = take(d.nd * 8);
  w.z = take(d.nd * 8);
  w.p = take(d.nd * 8);
  w.Ap = take(d.nd * 8);
  w.diag = take(d.nd * 8);
  w.total = o;
  return w;
}

bool valid_dims(int nelx, int nely, int nelz) {
  return nelx >= 1 && nely >= 1 && nelz >= 1 &&
         int64_t(nelx + 1) * (nely + 1) * (nelz + 1) * 3 < (int64_t(1) << 31);
}

// Grid3 for a call: material, pointers, reduction scratch carved from ws
// (counters zeroed by a kernel: the calls may be captured into graphs).
int make_grid3(const topt_grid3* g, const double* s, const uint32_t* fix, void* ws, hipStream_t st, Grid3& G,
               Dims3& d, Ws3& w) {
  if (g == nullptr) return fail(TOPT_EINVAL, "grid is NULL");
  if (!valid_dims(g->nelx, g->nely, g->nelz)) return fail(TOPT_EINVAL, "grid must have nelx, nely, nelz >= 1 "
                                                                        "and fewer than 2^31 DOFs");
  if (g->Ke == nullptr) return fail(TOPT_EINVAL, "grid Ke (device, 576 doubles) is NULL");
  std::memset(&G, 0, sizeof(G));
  G.nelx = g->nelx;
  G.nely = g->nely;
  G.nelz = g->nelz;
  G.Emin = g->E_min;
  G.penal = g->penal;
  G.Ke = g->Ke;
  G.s = s;
  G.fix = fix;
  d = dims3(g->nelx, g->nely, g->nelz);
  G.nbn = d.nbn;
  G.nbe = d.nbe;
  w = ws3(d);
  if (ws != nullptr) {
    char* b = (char*)ws;
    G.part = (double*)(b + w.part);
    G.ctr = (unsigned*)(b + w.ctr);
    G.cg = (Cg3*)(b + w.cg);
    hipLaunchKernelGGL(k3_zero_words, dim3(std::max<int64_t>(1, std::min<int64_t>(64, cdiv(w.ctr_words, 256)))),
                       dim3(256), 0, st, G.ctr, (long)w.ctr_words);
    LAUNCH_TRY();
  }
  return TOPT_OK;
}

// Filter stencil in the reference's table order (filter.py:85-94): jx outer,
// jy, jz inner over [-ceil(r), ceil(r)], kept when w = rmin - dist > 0
// (integer offsets from element centres), weight stored as float32.
std::vector<Tap3> build_taps3(double rmin) {
  std::vector<Tap3> taps;
  const int rc = (int)std::ceil(rmin);
  for (int dx = -rc; dx <= rc; ++dx)
    for (int dy = -rc; dy <= rc; ++dy)
      for (int dz = -rc; dz <= rc; ++dz) {
        const double w = rmin - std::sqrt((double)(dx * dx + dy * dy + dz * dz));
        if (w > 0.0) taps.push_back(Tap3{dx, dy, dz, 0, (double)(float)w});
      }
  return taps;
}

int upload_taps(double rmin, void* ws, const Ws3& w, hipStream_t st, const Tap3** dt, int* ntaps) {
  if (!(rmin > 0.0)) return fail(TOPT_EINVAL, "rmin must be > 0");
  if (std::ceil(rmin) > kMaxTapRadius) return fail(TOPT_EINVAL, "rmin too large for the 3-D filter (max 12)");
  const std::vector<Tap3> taps = build_taps3(rmin);
  Tap3* d = (Tap3*)((char*)ws + w.taps);
  HIP_TRY(hipMemcpyAsync(d, taps.data(), taps.size() * sizeof(Tap3), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));   // the taps come from pageable memory
  *dt = d;
  *ntaps = (int)taps.size();
  return TOPT_OK;
}

// Pinned slots for the solve's state snapshots: one pair per host thread,
// allocated on its first solve and kept until the process ends (not freed
// from a thread-exit destructor, which may run after the runtime is gone).
struct Poll {
  Cg3* host = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};
thread_local Poll g_poll;

int poll_slots(Poll& P) {
  if (!P.host && hipHostMalloc(&P.host, 2 * sizeof(Cg3), hipHostMallocDefault) != hipSuccess) {
    P.host = nullptr;
    return fail(TOPT_ENOMEM, "hipHostMalloc PCG status");
  }
  for (hipEvent_t& e : P.ev)
    if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return TOPT_OK;
}

}  // namespace

extern "C" {

int64_t topt3_workspace_bytes(int32_t nelx, int32_t nely, int32_t nelz) {
  if (!valid_dims(nelx, nely, nelz)) return -1;
  return ws3(dims3(nelx, nely, nelz)).total;
}

int topt3_scale(const topt_grid3* g, const double* rho, double* s, void* stream) {
  if (rho == nullptr || s == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  Grid3 G;
  Dims3 d;
  Ws3 w;
  int rc = make_grid3(g, nullptr, nullptr, nullptr, st, G, d, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k3_scale, dim3(d.nbe), dim3(TOPT_BLOCK), 0, st, G, rho, s);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt3_diagonal(const topt_grid3* g, const double* s, const uint32_t* fix, double* diag, void* stream) {
  if (s == nullptr || fix == nullptr || diag == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  Grid3 G;
  Dims3 d;
  Ws3 w;
  int rc = make_grid3(g, s, fix, nullptr, st, G, d, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k3_diag, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, diag);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt3_matvec(const topt_grid3* g, const double* p, double* Ap, const double* s, const uint32_t* fix,
                 double* pAp, void* ws, void* stream) {
  if (p == nullptr || Ap == nullptr || s == nullptr || fix == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  if (pAp != nullptr && ws == nullptr) return fail(TOPT_EINVAL, "p.Ap needs the workspace");
  hipStream_t st = (hipStream_t)stream;
  Grid3 G;
  Dims3 d;
  Ws3 w;
  int rc = make_grid3(g, s, fix, pAp ? ws : nullptr, st, G, d, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k3_matvec, dim3(d.nbt), dim3(TOPT_BLOCK), 0, st, G, p, Ap, pAp ? (int)MV3_DOT : (int)MV3_PLAIN,
                     pAp);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt3_elemblock_apply(const topt_grid3* g, const double* Ke_inv, double alpha_reg, const double* r, double* z,
                          const double* s, const uint32_t* fix, double* rz_rr, void* ws, void* stream) {
  if (Ke_inv == nullptr || r == nullptr || z == nullptr || s == nullptr || fix == nullptr || rz_rr == nullptr ||
      ws == nullptr)
    return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  Grid3 G;
  Dims3 d;
  Ws3 w;
  int rc = make_grid3(g, s, fix, ws, st, G, d, w);
  if (rc) return rc;
  G.Ki = Ke_inv;
  G.alpha_reg = alpha_reg;
  hipLaunchKernelGGL(k3_elemblock, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, r, z, (double*)nullptr, 0, rz_rr);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt3_sensitivity(const topt_grid3* g, const double* u, const double* rho, double* dc, double* dv,
                      double* compliance, void* ws, void* stream) {
  if (u == nullptr || rho == nullptr || dc == nullptr || compliance == nullptr || ws == nullptr)
    return fail(TOPT_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  Grid3 G;
  Dims3 d;
  Ws3 w;
  int rc = make_grid3(g, nullptr, nullptr, ws, st, G, d, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k3_sens, dim3(d.nbe), dim3(TOPT_BLOCK), 0, st, G, u, rho, dc, compliance);
  if (dv != nullptr)
    hipLaunchKernelGGL(k3_fill, dim3((unsigned)std::min<int64_t>(1024, d.nbe)), dim3(TOPT_BLOCK), 0, st, dv,
                       (long)d.ne, 1.0);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt3_filter(int32_t nelx, int32_t nely, int32_t nelz, double rmin, int32_t mode, const double* rho,
                 const double* dc, double* out, double* Hs, void* ws, void* stream) {
  if (out == nullptr || Hs == nullptr || ws == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  if (mode < 0 || mode > 2) return fail(TOPT_EINVAL, "3-D filter modes: density, dc average, sensitivity");
  if ((mode == 0 || mode == 2) && rho == nullptr) return fail(TOPT_EINVAL, "rho is NULL");
  if ((mode == 1 || mode == 2) && dc == nullptr) return fail(TOPT_EINVAL, "dc is NULL");
  if (!valid_dims(nelx, nely, nelz)) return fail(TOPT_EINVAL, "invalid 3-D grid");
  hipStream_t st = (hipStream_t)stream;
  const Dims3 d = dims3(nelx, nely, nelz);
  const Ws3 w = ws3(d);
  const Tap3* taps;
  int ntaps;
  int rc = upload_taps(rmin, ws, w, st, &taps, &ntaps);
  if (rc) return rc;
  hipLaunchKernelGGL(k3_filter, dim3(d.nbe), dim3(TOPT_BLOCK), 0, st, nelx, nely, nelz, taps, ntaps, (int)F_HS,
                     rho, dc, (const double*)nullptr, Hs);
  hipLaunchKernelGGL(k3_filter, dim3(d.nbe), dim3(TOPT_BLOCK), 0, st, nelx, nely, nelz, taps, ntaps, (int)mode,
                     rho, dc, (const double*)Hs, out);
  LAUNCH_TRY();
  return TOPT_OK;
}

int topt3_filter_tables(int32_t nelx, int32_t nely, int32_t nelz, double rmin, int32_t width, int32_t* nb,
                        float* wt, int32_t* cnt, void* ws, void* stream) {
  if (nb == nullptr || wt == nullptr || cnt == nullptr || ws == nullptr) return fail(TOPT_EINVAL, "NULL argument");
  if (!valid_dims(nelx, nely, nelz)) return fail(TOPT_EINVAL, "invalid 3-D grid");
  const Dims3 d = dims3(nelx, nely, nelz);
  const Ws3 w = ws3(d);
  hipStream_t st = (hipStream_t)stream;
  const Tap3* taps;
  int ntaps;
  int rc = upload_taps(rmin, ws, w, st, &taps, &ntaps);
  if (rc) return rc;
  if (ntaps > width) return fail(TOPT_EINVAL, "table width smaller than the stencil");
  hipLaunchKernelGGL(k3_filter_tables, dim3(d.nbe), dim3(TOPT_BLOCK), 0, st, nelx, nely, nelz, taps, ntaps, width,
                     nb, wt, cnt);
  LAUNCH_TRY();
  HIP_TRY(hipStreamSynchronize(st));
  return TOPT_OK;
}

int topt3_solve(const topt_grid3* g, int32_t precond, const double* Ke_inv, double alpha_reg, const double* f,
                const uint32_t* fix, const double* s, double* u, int32_t warm, double tol, int32_t max_iter,
                int32_t* n_iter, int32_t* status, void* ws, void* stream) {
  if (f == nullptr || fix == nullptr || s == nullptr || u == nullptr || ws == nullptr)
    return fail(TOPT_EINVAL, "NULL argument");
  if (precond != TOPT_PC_NONE && precond != TOPT_PC_JACOBI && precond != TOPT_PC_ELEMBLOCK)
    return fail(TOPT_EINVAL, "3-D preconditioners: none, jacobi, block_jacobi (element block)");
  if (precond == TOPT_PC_ELEMBLOCK && Ke_inv == nullptr) return fail(TOPT_EINVAL, "block_jacobi needs Ke_inv");
  if (max_iter < 0) return fail(TOPT_EINVAL, "max_iter must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  Grid3 G;
  Dims3 d;
  Ws3 w;
  int rc = make_grid3(g, s, fix, ws, st, G, d, w);
  if (rc) return rc;
  if ((rc = poll_slots(g_poll))) return rc;
  G.precond = precond;
  G.Ki = Ke_inv;
  G.alpha_reg = alpha_reg;
  char* b = (char*)ws;
  double *r = (double*)(b + w.r), *z = (double*)(b + w.z), *p = (double*)(b + w.p), *Ap = (double*)(b + w.Ap),
         *dg = (double*)(b + w.diag);
  const unsigned nbv = (unsigned)std::min<int64_t>(4096, cdiv(d.nd, TOPT_BLOCK));
  // setup (solver.py:55-104)
  if (precond == TOPT_PC_JACOBI) hipLaunchKernelGGL(k3_diag, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, dg);
  if (!warm) hipLaunchKernelGGL(k3_fill, dim3(nbv), dim3(TOPT_BLOCK), 0, st, u, (long)d.nd, 0.0);
  hipLaunchKernelGGL(k3_zero_fixed, dim3(nbv), dim3(TOPT_BLOCK), 0, st, u, fix, (long)d.nd);
  hipLaunchKernelGGL(k3_matvec, dim3(d.nbt), dim3(TOPT_BLOCK), 0, st, G, (const double*)u, Ap, (int)MV3_PLAIN,
                     (double*)nullptr);
  hipLaunchKernelGGL(k3_cg_init, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, f, (const double*)Ap, u, r, z, p,
                     (const double*)dg, tol, (int)max_iter);
  if (precond == TOPT_PC_ELEMBLOCK)
    hipLaunchKernelGGL(k3_elemblock, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, (const double*)r, z, p, 1,
                       (double*)nullptr);
  hipLaunchKernelGGL(k3_cg_zero_u, dim3(nbv), dim3(TOPT_BLOCK), 0, st, G, u, (long)d.nd);
  LAUNCH_TRY();
  // iterations in chunks; the state of chunk c is polled while chunk c+1 runs
  auto enqueue = [&](int count, int slot) -> int {
    for (int k = 0; k < count; ++k) {
      hipLaunchKernelGGL(k3_matvec, dim3(d.nbt), dim3(TOPT_BLOCK), 0, st, G, (const double*)p, Ap, (int)MV3_CG,
                         (double*)nullptr);
      hipLaunchKernelGGL(k3_cg_update, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, u, r, z, (const double*)p,
                         (const double*)Ap, (const double*)dg, (long)d.nd);
      if (precond == TOPT_PC_ELEMBLOCK)
        hipLaunchKernelGGL(k3_elemblock, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, (const double*)r, z,
                           (double*)nullptr, 2, (double*)nullptr);
      hipLaunchKernelGGL(k3_cg_pupdate, dim3(d.nbn), dim3(TOPT_BLOCK), 0, st, G, (const double*)z, p, (long)d.nd);
    }
    LAUNCH_TRY();
    HIP_TRY(hipMemcpyAsync(&g_poll.host[slot], G.cg, sizeof(Cg3), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(g_poll.ev[slot], st));
    return TOPT_OK;
  };
  int chunk = 8;
  int64_t issued = 0;
  for (int k = 0; k < 2; ++k) {
    if ((rc = enqueue(chunk, k))) return rc;
    issued += chunk;
  }
  int wait = 0;
  // a solve ends within max_iter iterations; the cap guards the host loop
  const int64_t cap = int64_t(max_iter) + 256;
  for (;;) {
    HIP_TRY(hipEventSynchronize(g_poll.ev[wait]));
    if (g_poll.host[wait].status != ST_RUN) break;
    if (issued > cap) return fail(TOPT_EINVAL, "3-D PCG loop did not terminate");
    chunk = std::min(chunk * 2, 32);
    if ((rc = enqueue(chunk, wait))) return rc;
    issued += chunk;
    wait ^= 1;
  }
  HIP_TRY(hipMemcpyAsync(&g_poll.host[wait], G.cg, sizeof(Cg3), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const Cg3& fin = g_poll.host[wait];
  if (n_iter) *n_iter = fin.status == ST_ZERO ? 0 : fin.it;
  if (status) *status = fin.status;
  return TOPT_OK;
}

}  // extern "C"
