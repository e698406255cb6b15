/*
 * chronos_amg.h — C ABI of the MI355X (gfx950) AMG solve phase.
 *
 * The AMG hierarchy (A_l, P_l, FSAI factors G_l, coarsest operator) is uploaded once; HIP
 * kernels then run the V-cycle and the Krylov loops.  The V-cycle runs in FP64 or in FP32
 * (CHR_FP32_CYCLE: A_l, P_l, G_l and the cycle vectors in FP32, coarsest solve in FP64)
 * inside an FP64 outer PCG.
 *
 * Conventions
 *   - Matrices: 0-based CSR, int32 row pointers and columns, columns sorted, FP64 values,
 *     host memory, borrowed for the duration of the call.
 *   - Dense vectors/blocks: FP64, column-major n x k with leading dimension ld (MATLAB
 *     layout).  `ptr_kind` says whether they are host (CHR_PTR_HOST) or device
 *     (CHR_PTR_DEVICE) pointers; device buffers must be ready when the call is made.
 *   - Every entry point returns 0 or a negative CHR_E* code; chr_last_error() returns a
 *     thread-local message.  No exception crosses the ABI.  The caller owns every buffer it
 *     passes; library objects are opaque and released by their *_destroy call.
 *   - A handle is bound to one HIP stream created by the library and is not thread safe;
 *     distinct handles may be used concurrently.  Calls return when their results are
 *     complete.
 *
 * Reference interfaces replaced (ChronosLab, paths relative to the reference root):
 *   chr_hier_create    the hierarchy struct list AMG_hrc built by
 *                      sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:82-130,288-302,347
 *                      (fields A, P, omega, nupre, nupost, smootherOp.left/right; L,S last)
 *   chr_vcycle_apply   M = @(x) AMG_Vcycle(AMG_prec, A, x)  (sources/utils/computeAMG.m:49,
 *                      sources/Preconditioner/AMG/aspAMG/AMG_Vcycle.m:1-54), called with an
 *                      n x k block as in sources/LinearSolver/PCG_CJ.m:38 and
 *                      sources/EigenSolver/lobpcg.m:216-218
 *   chr_pcg_solve      [x,iter,resvec,flag] = PCG_CJ(prodM,prodA,b,x0,itmax,tol)
 *                      (sources/LinearSolver/PCG_CJ.m:1-73)
 *   chr_spmv           prodA = @(x) A*x  (sources/Drivers/driver_AMG/aspAMG_main.m:200)
 */
#ifndef CHRONOS_AMG_H
#define CHRONOS_AMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CHR_OK = 0,
  CHR_EARG = -1,      /* bad argument */
  CHR_EALLOC = -2,    /* device or host allocation failure */
  CHR_EHIP = -3,      /* HIP runtime error */
  CHR_ENOTSPD = -4,   /* coarsest operator not positive definite */
  CHR_EBREAKDOWN = -5 /* Krylov breakdown (zero or non-finite p^T A p) */
};

enum { CHR_FP64 = 0, CHR_FP32_CYCLE = 1 };
enum { CHR_PTR_HOST = 0, CHR_PTR_DEVICE = 1 };

typedef struct chr_csr {
  int32_t nrows, ncols;
  const int32_t *rowptr;
  const int32_t *col;
  const double *val;
} chr_csr;

typedef struct chr_level {
  chr_csr A;       /* level operator; level 0: the operator the V-cycle uses */
  chr_csr P;       /* prolongation n_l x n_{l+1}; unused on the last level */
  chr_csr G;       /* FSAI factor (lower triangular, n_l x n_l); unused on the last level */
  double omega;    /* smoother damping: z = omega * G^T (G r) */
  int32_t nupre;   /* pre-smoothing sweeps */
  int32_t nupost;  /* post-smoothing sweeps */
} chr_level;

typedef struct chr_hier chr_hier;
typedef struct chr_mat chr_mat;

/* Per-hierarchy figures used for bandwidth accounting. */
typedef struct chr_hier_stats {
  int32_t nlev;
  int32_t precision;
  int64_t n[16];              /* rows per level */
  double vcycle_bytes;        /* bytes one V-cycle (k = 1) streams in the stored layouts */
  double vcycle_csr_bytes;    /* algorithmic bytes of one V-cycle: the same products with
                                 plain CSR (4-byte column indices, row pointers) */
  double level_bytes[16];     /* algorithmic bytes per level (a level's direct solves included) */
  double pcg_iter_bytes;      /* algorithmic: V-cycle + outer FP64 CSR product + 18 n_0 * 8 */
  double device_bytes;        /* device memory held by the hierarchy */
} chr_hier_stats;

const char *chr_last_error(void);

/* Select the HIP device used by subsequent calls on this thread. */
int chr_init(int32_t device);

/* Upload a hierarchy of nlev levels (the last one carries only A).  precision is
 * CHR_FP64 or CHR_FP32_CYCLE.  Transposes of P and G are formed on upload. */
int chr_hier_create(int32_t nlev, const chr_level *levels, int32_t precision, chr_hier **out);
int chr_hier_destroy(chr_hier *h);
int chr_hier_get_stats(const chr_hier *h, chr_hier_stats *st);

/* Layout the autotuner chose for the single-vector products of one level matrix (which =
 * 'A','G','T','P','R' = P^T, or a composite operator of a fused level: 'r' = [-P^T A | P^T],
 * 'u' = [A P | A], 'w' = [P | omega G^T G] or [P | omega G^T], 's' = omega G^T G; CHR_EARG
 * when the level has no such composite): per = nonzeros per thread of the CSR-stream kernel, minus that for
 * the packed-code stream kernel, or 0 for the sliced-ELL kernel; code_bits = bits per
 * stored column code (stream: 16 for block offsets, 32 for columns; packed: 16; sliced
 * ELL: 8 for dictionary codes, 16 for row-relative offsets, 32 for columns); nnz =
 * nonzeros; stored_bytes = bytes one product reads of the matrix in that layout.  Any
 * output may be NULL. */
int chr_hier_layout(const chr_hier *h, int32_t level, char which, int32_t *per,
                    int32_t *code_bits, int64_t *nnz, double *stored_bytes);

/* x = M b for an n x k block (k <= 16). */
int chr_vcycle_apply(chr_hier *h, int32_t k, const double *b, int64_t ldb, double *x,
                     int64_t ldx, int32_t ptr_kind);

/* Outer operator: a device copy of a CSR matrix in FP64. */
int chr_mat_create(const chr_csr *A, chr_mat **out);
int chr_mat_destroy(chr_mat *m);
/* y = A x (FP64). */
int chr_spmv(chr_mat *A, const double *x, double *y, int32_t ptr_kind);

/* PCG_CJ: preconditioner = V-cycle of h; prodA = A (NULL: the level-0 operator of h in
 * FP64).  x holds x0 on entry and the solution on exit.  resvec has room for itmax+1
 * entries (may be NULL).  flag = 0 converged, 1 not converged (PCG_CJ.m:123-125). */
int chr_pcg_solve(chr_hier *h, chr_mat *A, const double *b, double *x, int32_t itmax,
                  double tol, int32_t ptr_kind, int32_t *iter, double *resvec, int32_t *flag);

/* No counterpart in the reference.  For callers that cannot allocate device memory once they
 * run (serving loops, harnesses that forbid hipMalloc after start-up): chr_pcg_solve with
 * device pointers whose working memory - r, z, the two p buffers, Ap, the FP32 copy of r in
 * the FP32 cycle, the partial sums, the residual history and the solver state - lives in a
 * fixed device array of CHR_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until
 * the process ends.  The call makes no hipMalloc / hipFree and instantiates no graph: its
 * kernels are enqueued on the handle's stream one iteration at a time, with a wait for the
 * stop flag after each.  b, x: device pointers (x holds x0 on entry and is updated in
 * place); resvec (host, itmax+1 entries, may be NULL), iter and flag as for chr_pcg_solve,
 * whose results it equals bit for bit.  A solve needing more than the array (about 40 n
 * bytes, 44 n in the FP32 cycle, plus 8 (itmax+1) and the partial sums: n up to about 1.5
 * million) returns CHR_EARG before any work.  Call it on the device the hierarchy was created
 * on.  Every call on one GPU shares that GPU's array and leaves its contents behind; calls
 * that overlap on one GPU, through any handle, race. */
#define CHR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int chr_pcg_solve_noalloc(chr_hier *h, chr_mat *A, const double *b, double *x, int32_t itmax,
                          double tol, int32_t *iter, double *resvec, int32_t *flag);

/* lobpcg (sources/EigenSolver/lobpcg.m:21-467, B_orthonormalize.m:1-25) with prodB = [], as
 * sources/Drivers/driver_Eigen/EigDriver.m:202-226 calls it (the reference has no caller with
 * a non-empty prodB): prodA = A (NULL: the level-0 operator of h), prodM = V-cycle of h (h
 * NULL: no preconditioner, A required).  X: n x neig column-major (ldx), X0 on entry and the
 * eigenvectors on exit (neig <= 16).  Y: the constraint block blockY, n x m column-major
 * (ldy), m <= 48 (m = 0 or Y NULL: none): X0, the preconditioned residuals and the final X
 * are projected off it as lobpcg.m:64-75,228-230,384-387 do; CHR_EARG when Y^T Y is not
 * positive definite.  lambda[neig]; resnorm_hist and lambda_hist, when not
 * NULL, are (itmax+1) x neig column-major histories.  ierr: 0 converged, -1 not converged
 * (X = best iterate), 1 X0 or a Rayleigh-Ritz step rank deficient, 2 B-orthonormalisation
 * of the residual block failed (lobpcg.m:86-89,242-246,401-414). */
int chr_lobpcg(chr_hier *h, chr_mat *A, int32_t n, int32_t neig, double *X, int64_t ldx,
               const double *Y, int32_t m, int64_t ldy, int32_t largest, int32_t reslambda_check,
               int32_t itmax, double tol, int32_t restart_control, int32_t ptr_kind, int32_t *niter,
               double *lambda, double *resnorm_hist, double *lambda_hist, int32_t *ierr);

/* Adjoint gradient of a scalar loss for the cell-centred finite-volume diffusion operator
 * (harmonic face transmissibilities T_ij = 2 k_i k_j/(k_i+k_j) = -A_ij, Dirichlet terms
 * proportional to k_i): given the forward field x (A x = b) and the adjoint field lam
 * (A^T lam = A lam = dL/dx, e.g. from chr_pcg_solve on the same hierarchy), writes
 * g = dL/d(ln k) on A's sparsity pattern in one fused kernel.  A: NULL = the level-0
 * operator of h.  No reference counterpart (it serves the differentiable solve). */
int chr_fv_grad(chr_hier *h, chr_mat *A, const double *k, const double *x, const double *lam,
                double *g, int32_t ptr_kind);

/* Batched solves: nsys independent systems, each described as for chr_hier_create
 * (levels[s][0..nlev[s]-1]), merged into one block-diagonal hierarchy on the device.  The
 * merged V-cycle applied to a system's rows equals that system's own V-cycle bit for bit.
 * No reference counterpart (the reference solves one system per call of PCG_CJ). */
typedef struct chr_batch chr_batch;
int chr_batch_create(int32_t nsys, const int32_t *nlev, const chr_level *const *levels,
                     int32_t precision, chr_batch **out);
int chr_batch_destroy(chr_batch *bt);
/* The merged hierarchy (owned by the batch; e.g. for chr_hier_get_stats). */
chr_hier *chr_batch_hier(chr_batch *bt);
/* x = M b for every system: b and x hold the systems' blocks stacked in system order
 * (sum of n_s rows) x k, column-major with leading dimension = total rows (k <= 16). */
int chr_batch_vcycle(chr_batch *bt, int32_t k, const double *b, double *x, int32_t ptr_kind);
/* PCG_CJ (sources/LinearSolver/PCG_CJ.m:1-73) for every system and every column of its
 * right-hand-side block, all advancing together: one launch of a captured graph of 16
 * consecutive iterations (CHRONOS_BATCH_GRAPH_ITERS) solves the batch when every solve
 * converges within them, further launches only otherwise; each (system, column) keeps its
 * own scalars and stops at its own tolerance.  b, x as for chr_batch_vcycle (x holds x0).
 * iters and flags (nsys*k, index s*k + c) receive the iteration count and flag (0
 * converged, 1 not converged, -1 breakdown); resvec, when not NULL, (itmax+1) x (nsys*k)
 * column-major.  At most 1024 solves per call. */
int chr_batch_pcg(chr_batch *bt, int32_t k, const double *b, double *x, int32_t itmax,
                  double tol, int32_t ptr_kind, int32_t *iters, int32_t *flags, double *resvec);

/* Storage for a result whose size is known only at the end: called once with the entry
 * count, it sets *col and *val to caller-owned arrays of that many entries (or returns
 * non-zero to abort). */
typedef int (*chr_reserve_fn)(void *sink, int64_t nnz, int32_t **col, double **val);

/* Galerkin coarse operator A_c = P^T (A P) on the GPU (replaces `Anext = Pf'*(A*Pf)`,
 * sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:277): A n x n, P n x nc, both CSR with
 * sorted columns.  Writes rowptr_C[nc+1], obtains col/val storage through reserve, and
 * fills it.  Bit-identical to the host set-up's product (Gustavson rows, products in the
 * left factor's column order, rows sorted, exact zeros dropped).  Returns 0, a negative
 * code, or 1 when a row of the product has more than 4096 distinct columns (nothing is
 * written; the host forms such products). */
int chr_galerkin(const chr_csr *A, const chr_csr *P, int32_t *rowptr_C, void *sink,
                 chr_reserve_fn reserve);
/* chs_galerkin_fn adapter for chs_build (chronos_setup.h); ctx unused. */
int chr_setup_galerkin(void *ctx, int32_t n, const int32_t *rowptr_A, const int32_t *col_A,
                       const double *val_A, int32_t nc, const int32_t *rowptr_P,
                       const int32_t *col_P, const double *val_P, int32_t *rowptr_C,
                       void *sink, chr_reserve_fn reserve);

/* EMIN's CG operator for chs_build (the open / apply / close of chs_emin_kp_ops in
 * chronos_setup.h): y = (A X) restricted to the fine rows' interpolation pattern (rowptr_T,
 * col_T; columns ascending), each entry summed over A's row in order, bit for bit the host
 * product (KP_spmat.cpp as restated in the host library).  open uploads A and the pattern;
 * apply moves X (pattern order) in and y out; close frees the level.  No reference
 * counterpart (the reference runs this product on the CPU). */
int chr_setup_emin_open(void *ctx, int32_t n, int32_t nc, const int32_t *rowptr_A, const int32_t *col_A,
                        const double *val_A, const int32_t *rowptr_T, const int32_t *col_T, void **handle);
int chr_setup_emin_apply(void *handle, const double *x, double *y);
void chr_setup_emin_close(void *handle);

/* Adaptive symmetric FSAI factor of A computed on the GPU, every row's pattern growth and
 * dense solves in one batched launch (one wave per row).  Replaces the reference's native
 * setup kernel [nterm_G, iat_G, ja_G, coef_G, is_posdef] = SYM_aFSAI_compute(...)
 * (sources/Preconditioner/AMG/smoother/MEX_SYM_AFSAI/SYM_aFSAI_compute.cpp:49-149, cores
 * compute_local_fsai.cpp:6-116 and cpt_afsai_coef.cpp:316-554, called from afsai_cpp.m:50),
 * with that gateway's caller-allocated outputs: rowptr_G[n+1], and col_G / val_G of
 * n * (nstep*step_size + 1) entries, of which the first *nnz_G are written (rows in
 * ascending column order, diagonal last).  A: host CSR, square, 0-based, columns sorted.
 * nstep * step_size <= 63.  tau = 0 (no dropping), as afsai_cpp.m:46 sets it.  is_posdef
 * = 0 when a pattern block or a row's Schur complement was not positive (the reference's
 * is_posdef = false).  The factor equals chs_afsai's (chronos_setup.h) bit for bit.
 * Synchronous; uses a stream of its own on the current device. */
int chr_afsai(const chr_csr *A, int32_t nstep, int32_t step_size, double eps,
              int32_t *rowptr_G, int32_t *col_G, double *val_G, int64_t *nnz_G,
              int32_t *is_posdef);
/* The same with the signature of chs_afsai_fn (chronos_setup.h), so that the host setup
 * chs_build can take its aFSAI factors from the GPU: set chs_params.afsai_fn to it (ctx
 * unused).  Replaces the same reference gateway as chr_afsai. */
int chr_setup_afsai(void *ctx, int32_t n, const int32_t *rowptr, const int32_t *col,
                    const double *val, int32_t nstep, int32_t step_size, double eps,
                    int32_t *rowptr_G, int32_t *col_G, double *val_G, int64_t *nnz_G,
                    int32_t *is_posdef);

/* Smoother damping omega = min(1, 1.9 / lambda_max(G A G^T)) with lambda_max by Lanczos
 * (40 steps at most, full reorthogonalisation, stop at 1e-7 relative change) on the GPU.
 * Replaces the eigs call of sources/Preconditioner/AMG/smoother/smoother.m:72-81; the
 * start vector is U[-1,1) from the shared counter RNG (seed, stream 7) as in chs_fsai_omega
 * (chronos_setup.h), whose lambda it matches to rounding. */
int chr_fsai_omega(const chr_csr *A, const chr_csr *G, uint64_t seed, double *omega,
                   double *lambda);
/* The same with the signature of chs_omega_fn (chronos_setup.h), for chs_params.omega_fn. */
int chr_setup_omega(void *ctx, int32_t n, const int32_t *rowptr_A, const int32_t *col_A,
                    const double *val_A, const int32_t *rowptr_G, const int32_t *col_G,
                    const double *val_G, uint64_t seed, double *omega, double *lambda);

/* Timing helpers (HIP events on the handle's stream, device-resident data):
 * mean milliseconds of one V-cycle (k = 1) over reps cycles, captured back to back in a
 * graph of at most 10 cycles (as the solver's iteration graph holds its cycle) that is
 * launched until reps cycles have run, after one warm-up launch; the FP32 cycle reads an
 * FP32 copy of its input in the first smoother product, as the solver does. */
int chr_time_vcycle(chr_hier *h, int32_t reps, double *ms);
/* Mean milliseconds of one launch of a level SpMV: which = 'A', 'G', 'T' (G^T), 'P',
 * 'R' (P^T), each launch timed on its own after a 512 MiB device copy has flushed the
 * 256 MB Infinity Cache (as in a V-cycle, where the other products evict the matrix).
 * bytes receives the algorithmic bytes of one launch: the matrix as plain CSR (4-byte
 * column indices and row pointers), the input vector read once and the output written
 * once. */
int chr_time_spmv(chr_hier *h, int32_t level, char which, int32_t reps, double *ms,
                  double *bytes);

#ifdef __cplusplus
}
#endif
#endif
