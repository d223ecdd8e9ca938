/* libtopt: SIMP topology-optimization hot path for AMD Instinct MI355X (gfx950).
 *
 * C ABI only: plain device pointers, sizes and host scalars; every call is
 * asynchronous on the given HIP stream (passed as void*) unless stated.
 * Functions return 0 on success or a negative TOPT_E* code; topt_last_error()
 * returns a thread-local message.
 *
 * Each entry point replaces a Taichi kernel or host driver of the reference
 * (tomohiron907/toptimization, paths relative to src/toptimization/).
 *
 * Layout: node n = ix*(nely+1)+iy with DOFs interleaved (x, y) -> 2 scalars
 * per node; element e = ex*nely+ey; fixed DOFs as a bitmask (bit 2n+c of a
 * uint32 array of ceil(2*nodes/32) words).
 */
#ifndef TOPT_H
#define TOPT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision of the PCG vectors */
#define TOPT_F32 0   /* mixed: FP32 matvec/vectors, FP64 dots + FP64 residual refresh */
#define TOPT_F64 1   /* plain FP64 */

/* preconditioners (reference solver.py:12-15; BLOCK2 is new) */
#define TOPT_PC_NONE 0
#define TOPT_PC_JACOBI 1
#define TOPT_PC_ELEMBLOCK 2   /* reference "block_jacobi" (preconditioner.py) */
#define TOPT_PC_BLOCK2 3      /* nodal 2x2 block Jacobi */
#define TOPT_PC_MG 4          /* geometric multigrid V(1,1) cycle, Galerkin coarse levels */

/* filter modes (reference optimizer/filter.py) */
#define TOPT_FILTER_DENSITY 0     /* rhof = H rho / Hs          filter.py:106-124 */
#define TOPT_FILTER_DCAVG 1       /* dcf  = H dc / Hs           filter.py:127-149 */
#define TOPT_FILTER_SENS 2        /* dcf  = H(rho dc)/(rho Hs)  filter.py:152-171 */
#define TOPT_FILTER_TRANSPOSE 3   /* dcf  = H^T (dc / Hs)       exact chain rule (new) */

/* PCG status */
#define TOPT_ST_RUN 0
#define TOPT_ST_CONV 1
#define TOPT_ST_BREAK 2    /* |pAp| < 1e-30 (solver.py:114-115): reports max_iter */
#define TOPT_ST_MAXIT 3
#define TOPT_ST_ZERO 4     /* |f| < 1e-30 (solver.py:60-63): u = 0, 0 iterations */
#define TOPT_ST_DONE 5     /* topt_plan_advance: the design's OC loop has stopped */

/* error codes */
#define TOPT_OK 0
#define TOPT_EINVAL -1
#define TOPT_EHIP -2
#define TOPT_ENOMEM -3

/* Per-problem workspace buffers (byte offsets from topt_problem_layout). */
enum {
  TOPT_BUF_U = 0,   /* double[2*nodes]   displacement (FP64 solution)             */
  TOPT_BUF_F,       /* double[2*nodes]   load vector (caller writes)              */
  TOPT_BUF_FIX,     /* uint32[words]     fixed-DOF bitmask (caller writes)        */
  TOPT_BUF_RHO,     /* double[elems]     design density (caller writes initial)   */
  TOPT_BUF_RHOF,    /* double[elems]     filtered (physical) density              */
  TOPT_BUF_DC,      /* double[elems]     compliance sensitivity                   */
  TOPT_BUF_DCF,     /* double[elems]     filtered sensitivity                     */
  TOPT_BUF_HS,      /* double[elems]     filter row sums                          */
  TOPT_BUF_X, TOPT_BUF_R, TOPT_BUF_Z, TOPT_BUF_P0, TOPT_BUF_P1, TOPT_BUF_AP,
  TOPT_BUF_PRE, TOPT_BUF_PRE2, TOPT_BUF_S, TOPT_BUF_S64,
  TOPT_BUF_TAPS, TOPT_BUF_CG, TOPT_BUF_OC, TOPT_BUF_PART, TOPT_BUF_OPART, TOPT_BUF_CTR,
  TOPT_BUF_MG,      /* multigrid levels (TOPT_PC_MG only): vectors, element matrices, coarse inverse */
  TOPT_BUF_COUNT
};

/* One design problem: 2-D Q4 plane-stress grid, SIMP parameters, workspace. */
typedef struct topt_problem {
  int32_t nelx, nely;
  int32_t precond;       /* TOPT_PC_* */
  int32_t max_cg_iter;   /* solver.max_cg_iterations (problem.py:117) */
  double E_min, penal;   /* material / SIMP (problem.py:103-108) */
  double cg_tol;         /* relative residual tolerance (solver.py:45, 65) */
  double volfrac, move, rmin;
  double fnorm2;         /* |f|^2 of the load the caller wrote into TOPT_BUF_F */
  double delta;          /* mixed path: refresh residual when |r| drops by this factor (0.1) */
  double Ke[64];         /* element stiffness, row-major, E folded in (material.py:20-86) */
  void* ws;              /* device workspace of topt_problem_layout() bytes */
  double oc_tol;         /* OC stop: it > 5 and change < oc_tol, or it == oc_max_iter */
  int32_t oc_max_iter;   /* (optimizer/simp.py:217; used by topt_plan_advance) */
  int32_t pad;
} topt_problem;

typedef struct topt_plan topt_plan;

/* Per-problem results of the last SIMP iteration (optimizer/simp.py:192-197). */
typedef struct topt_iter_stats {
  double compliance;   /* sum_e s_e u_e^T Ke u_e            (kernels.py:234-272) */
  double volume;       /* mean rho after the OC update      (kernels.py:330-336) */
  double change;       /* max |rho_new - rho|                (oc.py:46, 80) */
  double lambda;       /* last bisection midpoint            (oc.py:72) */
  int32_t n_cg;        /* PCG iterations, reference count semantics (solver.py:149,159) */
  int32_t cg_status;   /* TOPT_ST_* */
  int32_t oc_steps;    /* bisection steps taken (<= 50)      (oc.py:71) */
  int32_t refines;     /* FP64 residual refreshes (mixed path) */
  int32_t iter;        /* OC iterations completed */
  int32_t done;        /* the OC stopping rule has been met (simp.py:217) */
} topt_iter_stats;

const char* topt_last_error(void);
int topt_device_count(int* n);

/* ---- per-problem workspace (replaces fields.allocate, fem/fields.py:53-156) ---- */
/* Byte offsets of every TOPT_BUF_* buffer (256-byte aligned) and the total. */
int topt_problem_layout(int32_t nelx, int32_t nely, int32_t dtype, int32_t precond, double rmin,
                        int64_t* offsets, int64_t* total_bytes);
/* Build the filter stencil and row sums, zero u, reset the PCG/OC state.
 * Caller has written F, FIX and RHO. Synchronous on the stream. */
int topt_problem_init(const topt_problem* prob, int32_t dtype, void* stream);

/* ---- batched engine: one launch per kernel advances every problem ---- */
int topt_plan_create(const topt_problem* probs, int32_t nprob, int32_t dtype, topt_plan** out);
int topt_plan_destroy(topt_plan* plan);

/* Asynchronous SIMP driver (continuous batching at the iteration level).
 * Every problem cycles through its own OC iterations — [density filter ->
 * PCG setup] -> PCG -> [sensitivity -> dc filter -> OC] — independently of
 * the others: PCG launches cover the problems currently solving, and a
 * problem whose solve ends gets its OC step and its next solve set up at
 * the next chunk boundary while the others keep solving. Per problem:
 *   phase[i] (in/out): 0 = about to start an OC iteration, 1 = solving,
 *                      2 = solve ended, OC step pending, 3 = done
 *                      (stopping rule met);
 *   stats[i] (out): statistics of its latest completed OC iteration;
 *   done_now[i] (out): OC iterations completed during this call;
 *   cg_now[i] (out): PCG iterations of those OC iterations.
 * A problem's OC loop stops on the device (simp.py:217 with its oc_tol and
 * oc_max_iter). Returns when a problem becomes done, after max_chunks
 * chunks of 16 CG iterations, or when all are done; every problem is then at
 * a chunk boundary, so the caller may rebuild the plan with other problems
 * and call again (a solving problem resumes where it stopped; phase is the
 * only state the caller carries). Results are bitwise those of
 * topt_plan_iteration; the filtered density of a finished design is the
 * filter of its final rho. Replaces the host loop of simp.py:141-219 for a
 * batch; no reference counterpart. Blocks. Graph replay must be off. */
int topt_plan_advance(topt_plan* plan, int32_t sens_filter, int32_t max_chunks, int32_t* phase,
                      topt_iter_stats* stats, int32_t* done_now, int64_t* cg_now, void* stream);

/* Replay the PCG slot launches of every full chunk (32 CG iterations) from a
 * HIP graph captured once per plan (on a private stream, kernels only),
 * instead of enqueueing them one by one. Results are identical either way.
 * Replaces the per-iteration host loop of fem/solver.py:107-156 further; no
 * direct reference counterpart. Default off. */
int topt_plan_set_graph(topt_plan* plan, int32_t enable);
/* Density / sensitivity filter of every problem (filter.py:106-171). */
int topt_plan_filter(topt_plan* plan, int32_t mode, void* stream);
/* PCG solve K(rho_phys) u = f for every problem (fem/solver.py:27-159).
 * warm: 0 cold, 1 warm start from u, 2 warm iff the problem finished an OC
 * iteration (simp.py:160). use_filtered: rho_phys = RHOF (density filter) else RHO.
 * Blocks until all problems finished; writes n_iter[i] and status[i]. */
int topt_plan_solve(topt_plan* plan, int32_t warm, int32_t use_filtered, int32_t* n_iter,
                    int32_t* status, void* stream);
/* Compliance + sensitivities of every problem (fem/kernels.py:234-272). */
int topt_plan_sensitivity(topt_plan* plan, int32_t use_filtered, void* stream);
/* OC bisection + commit for every problem (optimizer/oc.py:52-93). */
int topt_plan_oc(topt_plan* plan, void* stream);
/* One SIMP iteration for every problem (optimizer/simp.py:141-197):
 * density mode (sens_filter=0) or sensitivity mode (1). The launches before
 * the PCG (density filter, solve setup) and after it (sensitivity, dc filter,
 * OC, statistics) are replayed from two HIP graphs per mode, captured on the
 * first call. Blocks when stats != NULL (fills stats[nprob]). */
int topt_plan_iteration(topt_plan* plan, int32_t sens_filter, topt_iter_stats* stats, void* stream);
/* Benchmark support: while enabled, every topt_plan_iteration call adds the
 * device time (HIP events on its stream) of its three phases to running sums:
 * ms3[0] density filter + solve setup, ms3[1] the PCG (all chunks of CG
 * iterations, including the chunk in flight when the stop is seen), ms3[2]
 * sensitivity, dc filter, OC bisection and statistics. Enabling resets the
 * sums. No reference counterpart. */
int topt_plan_set_timing(topt_plan* plan, int32_t enable);
int topt_plan_timing(topt_plan* plan, double* ms3, int64_t* calls);
/* Read back per-problem stats of the last iteration (blocks). */
int topt_plan_stats(topt_plan* plan, topt_iter_stats* stats, void* stream);

/* Benchmark support: run `reps` back-to-back launches of PCG slot kernel A
 * (which must be 0) as it runs inside a solve: the deferred u += alpha p of
 * the previous step (here with alpha = 0: u keeps its values), p = z + beta p
 * over the tile halo, Ap = K p, p.Ap; on the plan's current vectors. Returns
 * the elapsed GPU time (hipEvents on the stream) in *ms. Leaves every
 * problem marked converged. No reference counterpart. Blocks. */
int topt_plan_time_slot(topt_plan* plan, int32_t which, int32_t reps, float* ms, void* stream);

/* ---- op-level kernels (single problem, explicit buffers) ---- */
typedef struct topt_grid {
  int32_t nelx, nely;
  double E_min, penal;
  double Ke[64];
} topt_grid;

/* Workspace bytes for the op-level reductions of a grid. */
int64_t topt_op_workspace_bytes(int32_t nelx, int32_t nely);
/* s_e = E_min + rho_e^p (1 - E_min)  (fem/kernels.py:92). s in dtype. */
int topt_scale(const topt_grid* g, int32_t dtype, const double* rho, void* s, void* stream);
/* Ap = K(s) p, fixed rows identity (fem/kernels.py:143-176); if pAp != NULL
 * also p.Ap (FP64, device) (fem/fused_kernels.py:21-56). */
int topt_matvec(const topt_grid* g, int32_t dtype, const void* p, void* Ap, const void* s,
                const uint32_t* fix, double* pAp, void* ws, void* stream);
/* diag K(s), fixed -> 1, clamp 1e-20 (fem/kernels.py:76-101). */
int topt_diagonal(const topt_grid* g, int32_t dtype, const void* s, const uint32_t* fix,
                  void* diag, void* stream);
/* Nodal 2x2 block-Jacobi inverse: inv_diag[2*nodes] (xx, yy), inv_off[nodes] (xy). */
int topt_block2_inverse(const topt_grid* g, int32_t dtype, const void* s, const uint32_t* fix,
                        void* inv_diag, void* inv_off, void* stream);
/* Element-block Jacobi z = M r, r.z and r.r into rz_rr[2] (preconditioner.py:47-108). */
int topt_elemblock_apply(const topt_grid* g, int32_t dtype, const void* r, void* z,
                         const void* s, const uint32_t* fix, double* rz_rr, void* ws,
                         void* stream);
/* Compliance + dc (FP64) (fem/kernels.py:234-272); dv filled with 1 if non-NULL. */
int topt_sensitivity(const topt_grid* g, const double* u, const double* rho, double* dc,
                     double* dv, double* compliance, void* ws, void* stream);
/* Stencil filter (modes above). Hs: double[elems] scratch filled by this call. */
int topt_filter(int32_t nelx, int32_t nely, double rmin, int32_t mode, const double* rho,
                const double* dc, double* out, double* Hs, void* ws, void* stream);
/* Expand the stencil into the reference's padded tables (filter.py:52-74):
 * nb int32[elems*width] (-1 pad), wt float[elems*width] (0 pad), cnt int32[elems]. */
int topt_filter_tables(int32_t nelx, int32_t nely, double rmin, int32_t width, int32_t* nb,
                       float* wt, int32_t* cnt, void* ws, void* stream);
/* OC bisection + commit on rho in place (optimizer/oc.py:52-93). dv may be NULL (= 1).
 * out[4] (device): change, volume, last lambda, steps. */
int topt_oc_update(int32_t nelx, int32_t nely, double* rho, const double* dcf, const double* dv,
                   double volfrac, double move, double* out, void* ws, void* stream);
/* Allocation-free variant of topt_oc_update (no reference counterpart): for
 * callers that cannot allocate device memory or pass a workspace, such as
 * code under HIP graph capture and tight serving loops. Same arguments
 * without ws, same results bit for bit. Its scratch (OC state, reduction
 * counters, partial sums) is a static device array of
 * TOPT_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; grids
 * needing more (beyond about 32 million elements) are rejected with
 * TOPT_EINVAL. Makes no hipMalloc/hipFree/hipMemset call; the state is
 * cleared by a kernel at every call. All calls on one GPU share the array:
 * calls that overlap on one GPU (other streams, other threads) race. */
#define TOPT_NOALLOC_SCRATCH_BYTES (4 << 20)
int topt_oc_update_noalloc(int32_t nelx, int32_t nely, double* rho, const double* dcf, const double* dv,
                           double volfrac, double move, double* out, void* stream);
/* mean of an element field (fem/kernels.py:330-336) into out[0] (device). */
int topt_volume(int32_t nelx, int32_t nely, const double* v, double* out, void* ws, void* stream);

/* ---- 3-D H8 path (FP64; trilinear hexahedra on the structured grid) ----
 * Layout: node n = (ix*(nely+1) + iy)*(nelz+1) + iz with three DOFs per node
 * interleaved (x, y, z); element e = (ex*nely + ey)*nelz + ez (mesh.py:68-112);
 * fixed DOFs as a bitmask (bit 3n+c). Ke is a device array of 576 doubles,
 * the 24x24 H8 stiffness in the reference's local DOF order (material.py:89-171).
 * The OC update of a 3-D design is topt_oc_update with nely*nelz in place of
 * nely (it only reads element arrays). */
typedef struct topt_grid3 {
  int32_t nelx, nely, nelz, pad;
  double E_min, penal;
  const double* Ke;
} topt_grid3;

/* Workspace bytes of the 3-D calls (reductions, PCG vectors, filter taps). */
int64_t topt3_workspace_bytes(int32_t nelx, int32_t nely, int32_t nelz);
/* s_e = E_min + rho_e^p (1 - E_min) (fem/kernels.py:200). */
int topt3_scale(const topt_grid3* g, const double* rho, double* s, void* stream);
/* diag K(s), fixed -> 1, clamp 1e-20 (fem/kernels.py:104-129). */
int topt3_diagonal(const topt_grid3* g, const double* s, const uint32_t* fix, double* diag, void* stream);
/* Ap = K(s) p, fixed rows identity (fem/kernels.py:179-212); if pAp != NULL
 * also p.Ap (fem/fused_kernels.py:59-94). */
int topt3_matvec(const topt_grid3* g, const double* p, double* Ap, const double* s, const uint32_t* fix,
                 double* pAp, void* ws, void* stream);
/* Element-block Jacobi z = M r with r.z, r.r into rz_rr[2]
 * (fem/preconditioner.py:111-156); Ke_inv = inv(Ke + alpha_reg I) (device, 576). */
int topt3_elemblock_apply(const topt_grid3* g, const double* Ke_inv, double alpha_reg, const double* r, double* z,
                          const double* s, const uint32_t* fix, double* rz_rr, void* ws, void* stream);
/* Compliance and dc (fem/kernels.py:275-308); dv filled with 1 if non-NULL. */
int topt3_sensitivity(const topt_grid3* g, const double* u, const double* rho, double* dc, double* dv,
                      double* compliance, void* ws, void* stream);
/* 3-D filter over the reference's neighbourhood (optimizer/filter.py:75-97,
 * 106-171): modes TOPT_FILTER_DENSITY, _DCAVG, _SENS; Hs is filled. rmin <= 12. */
int topt3_filter(int32_t nelx, int32_t nely, int32_t nelz, double rmin, int32_t mode, const double* rho,
                 const double* dc, double* out, double* Hs, void* ws, void* stream);
/* The reference's padded 3-D tables (optimizer/filter.py:75-97). Blocks. */
int topt3_filter_tables(int32_t nelx, int32_t nely, int32_t nelz, double rmin, int32_t width, int32_t* nb,
                        float* wt, int32_t* cnt, void* ws, void* stream);
/* PCG K(s) u = f (fem/solver.py:27-167) with preconditioner TOPT_PC_NONE,
 * _JACOBI or _ELEMBLOCK, on the device (alpha, beta and the stopping tests in
 * device state); warm = 1 starts from u, else from 0; u[fixed] = 0. Blocks;
 * n_iter follows the reference's count (max_iter on breakdown or the cap, 0
 * for a zero load, which leaves u = 0). */
int topt3_solve(const topt_grid3* g, int32_t precond, const double* Ke_inv, double alpha_reg, const double* f,
                const uint32_t* fix, const double* s, double* u, int32_t warm, double tol, int32_t max_iter,
                int32_t* n_iter, int32_t* status, void* ws, void* stream);

/* Vector ops of the drop-in API, op codes (buffers by role x, y, z, w, v):
 *  0 FILL      x = a                              fem/kernels.py:19-23
 *  1 COPY      y = x                              fem/kernels.py:26-30
 *  2 AXPY      z = a x + y                        fem/kernels.py:33-37
 *  3 SCALE     z = a x                            fem/kernels.py:40-44
 *  4 DIV       z = x / y                          fem/kernels.py:65-73
 *  5 DOT       red[0] = x.y (FP64)                fem/kernels.py:47-62
 *  6 MAXDIFF   red[0] = max|x - y|                fem/kernels.py:339-345
 *  7 SUM       red[0] = sum x                     fem/kernels.py:330-336
 *  8 UPD_JAC   x=u += a w(=p); y=r -= a v(=Ap); z = r / d(=diag);
 *              red = (r.z, r.r)                   fem/fused_kernels.py:110-139
 *  9 UPD_NONE  as 8 with z = r                    fem/fused_kernels.py:142-166
 * 10 UPD_UR    x=u += a w; y=r -= a v             fem/fused_kernels.py:169-180
 * 11 UPD_P     w=p = z + a w                      fem/fused_kernels.py:183-187
 * 12 ZERO_FIXED x[i] = 0 where bit i of fix set   fem/solver.py:162-167 */
int64_t topt_blas1_workspace_bytes(void);
int topt_blas1(int32_t op, int32_t dtype, int64_t n, double a, void* x, void* y, void* z, void* w, void* v,
               const void* d, const uint32_t* fix, double* red, void* ws, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TOPT_H */
