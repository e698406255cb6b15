/*
 * chronos_setup.h — host-side (CPU, C++/OpenMP) producer of the AMG hierarchy that the
 * GPU solve phase consumes, plus the synthetic problem generators and the counter-based
 * RNG shared by every generator in this project.
 *
 * All matrices are 0-based CSR with int32 row pointers and int32 column indices, columns
 * sorted ascending inside every row, FP64 values.  The caller owns every buffer it passes
 * in; objects created by the library are opaque handles released by a matching *_free.
 * Every entry point returns 0 on success or a negative CHS_E* code; chs_last_error()
 * returns a thread-local message for the last failure.
 *
 * Reference counterparts (ChronosLab, paths relative to the reference root):
 *   chs_build                 sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:68-351
 *   chs_soc_cla               sources/Preconditioner/AMG/coarsen/cpt_SoC_cla.m:1-90
 *   chs_pmis                  sources/Preconditioner/AMG/coarsen/cpt_PMIS.m:1-102
 *                             + mark_isolated_nodes.m:1-8
 *   chs_prol_classical        sources/Preconditioner/AMG/prolong/MEX_CLAS_prol/
 *                             Classical_prolongation.cpp:40-147, ProlStripe_Classical.cpp:19-175
 *   chs_prol_exti             sources/Preconditioner/AMG/prolong/MEX_EXTI_Prol/
 *                             EXTI_prolongation.cpp:52-199, ONELP_ProlStripe_EXTI.cpp:27-293
 *                             (called by cpt_ProlMEX_EXTI.m:1-38 from prolong.m:84-87)
 *   chs_afsai                 sources/Preconditioner/AMG/smoother/MEX_SYM_AFSAI/
 *                             compute_local_fsai.cpp:6-116, cpt_afsai_coef.cpp:316-554
 *   chs_fsai_omega            sources/Preconditioner/AMG/smoother/smoother.m:72-81
 *   chs_galerkin              sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:277
 *   chs_bc_fix                sources/utils/computeAMG.m:5-10
 */
#ifndef CHRONOS_SETUP_H
#define CHRONOS_SETUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CHS_OK = 0,
  CHS_EARG = -1,     /* bad argument */
  CHS_EALLOC = -2,   /* allocation failure */
  CHS_ENOTSPD = -4,  /* FSAI found a non positive definite block (reference: is_posdef=false) */
  CHS_ECONSIST = -6  /* F/C count inconsistency (reference: cpt_aspAMG.m:212-215) */
};

/* Opaque CSR matrix owned by the library. */
typedef struct chs_mat chs_mat;
/* Opaque hierarchy owned by the library. */
typedef struct chs_hier chs_hier;

/* Replacement of the aFSAI factor computation inside chs_build (e.g. the GPU kernel of
 * libchronos_amg, chr_setup_afsai in chronos_amg.h).  Contract of chr_afsai: A in CSR,
 * outputs rowptr_G[n+1] and col_G / val_G of n * (nstep*step_size + 1) entries, nnz_G the
 * entries written; returns 0 or a negative code. */
typedef int (*chs_afsai_fn)(void *ctx, int32_t n, const int32_t *rowptr, const int32_t *col,
                            const double *val, int32_t nstep, int32_t step_size, double eps,
                            int32_t *rowptr_G, int32_t *col_G, double *val_G, int64_t *nnz_G,
                            int32_t *is_posdef);

/* Replacement of the smoother damping computation inside chs_build (e.g. chr_setup_omega
 * of libchronos_amg): omega = min(1, 1.9 / lambda_max(G A G^T)) for the level's A and G. */
typedef int (*chs_omega_fn)(void *ctx, int32_t n, const int32_t *rowptr_A, const int32_t *col_A,
                            const double *val_A, const int32_t *rowptr_G, const int32_t *col_G,
                            const double *val_G, uint64_t seed, double *omega, double *lambda);

/* Setup parameters: the fields of the reference's parm_AMG / parm_COARSEN / parm_SMOOTH
 * files that the classical + aFSAI path reads (Tests/parm_*, examples/AMG/parm_*). */
typedef struct chs_params {
  int32_t nlev_max;     /* parm_AMG line 1  (nLevMax, 10) */
  int32_t max_coarse;   /* parm_AMG line 2  (maxCoarseSZ, 1000) */
  double soc_tau;       /* parm_COARSEN line 2 (CLA tau) */
  int32_t nstep;        /* parm_SMOOTH line 4 (15) */
  int32_t step_size;    /* parm_SMOOTH line 5 (1) */
  double fsai_eps;      /* parm_SMOOTH line 6 (1e-3) */
  int32_t nupre;        /* parm_SMOOTH line 2 */
  int32_t nupost;       /* parm_SMOOTH line 3 */
  int32_t smoother;     /* 0 = afsai_sym, 1 = jacobi */
  int32_t prolong;      /* parm_PROLONG proltype: 0 = CLAS, 1 = EXTI */
  int32_t nthreads;     /* OpenMP threads (<=0: runtime default) */
  uint64_t seed;        /* PMIS tie-break stream and Lanczos start vector */
  int32_t verbose;
  chs_afsai_fn afsai_fn; /* NULL: chs_afsai on the host */
  void *afsai_ctx;       /* passed to afsai_fn */
  chs_omega_fn omega_fn; /* NULL: chs_fsai_omega on the host */
  void *omega_ctx;       /* passed to omega_fn */
} chs_params;

void chs_params_default(chs_params *p);
const char *chs_last_error(void);

/* ---- opaque matrix access -------------------------------------------------------- */
int chs_mat_from_csr(int32_t nrows, int32_t ncols, const int32_t *rowptr,
                     const int32_t *col, const double *val, chs_mat **out);
int chs_mat_dims(const chs_mat *m, int32_t *nrows, int32_t *ncols, int64_t *nnz);
int chs_mat_copy(const chs_mat *m, int32_t *rowptr, int32_t *col, double *val);
void chs_mat_free(chs_mat *m);

/* ---- counter-based RNG (SplitMix64 finaliser over a Weyl counter) ----------------- */
/* u64 number k of stream s under seed: mix64(mix64(seed ^ s*C1) + (k+1)*GAMMA). */
uint64_t chs_rng_u64(uint64_t seed, uint64_t stream, uint64_t k);
/* out[i] = lo + (hi-lo) * U[0,1) draw number offset+i (53-bit mantissa). */
void chs_rng_uniform(uint64_t seed, uint64_t stream, int64_t offset, int64_t n,
                     double lo, double hi, double *out);
/* Standard normals by Box-Muller on consecutive uniform pairs (2i, 2i+1). */
void chs_rng_normal(uint64_t seed, uint64_t stream, int64_t offset, int64_t n,
                    double *out);

/* ---- synthetic problems ------------------------------------------------------------ */
/* 7-point Laplacian (6 on the diagonal, -1 off) on an N^3 grid, Dirichlet boundary
 * eliminated, lexicographic order i = x + N*(y + N*z). */
int chs_gen_poisson7(int32_t N, chs_mat **out);
/* Cell-centred finite-volume diffusion on N^3 cells with per-cell permeability k (length
 * N^3, lexicographic), harmonic face transmissibilities 2 k_i k_j / (k_i + k_j), Dirichlet
 * half-cell terms 2 k_i on the x=0 and x=1 faces, no flux elsewhere. */
int chs_gen_diffusion(int32_t N, const double *k, chs_mat **out);
/* ln k ~ N(0, sigma^2) i.i.d. from the RNG stream (seed, 1). */
void chs_gen_lnk(int32_t N, double sigma, uint64_t seed, double *lnk);

/* ---- setup components ------------------------------------------------------------- */
/* Classical strength of connection.  strong[e] = +1 strong / -1 weak for every stored
 * entry e of A (diagonal forced to +1), i.e. the sign pattern S_patt aligned with A. */
int chs_soc_cla(const chs_mat *A, double tau, int8_t *strong);
/* Dirichlet marking + sequential PMIS.  rand01[i] is the tie-break added to node i's
 * weight.  Output fcnode[i] = c >= 0 (coarse index) or -1 (fine); returns nc in *nc. */
int chs_pmis(const chs_mat *A, const int8_t *strong, const double *rand01,
             int32_t *fcnode, int32_t *nc);
/* Classical direct interpolation with strong-fine correction. */
int chs_prol_classical(const chs_mat *A, const int8_t *strong, const int32_t *fcnode,
                       int32_t nc, int nthreads, chs_mat **P);
/* Extended+i distance-two interpolation (the reference's EXTI prolongation). */
int chs_prol_exti(const chs_mat *A, const int8_t *strong, const int32_t *fcnode, int32_t nc,
                  int nthreads, chs_mat **P);
/* Adaptive symmetric FSAI (lower-triangular G, diagonal last in every row). */
int chs_afsai(const chs_mat *A, int32_t nstep, int32_t step_size, double eps,
              int nthreads, chs_mat **G, int32_t *is_posdef);
/* omega = min(1, 1.9 / lambda_max(G A G^T)) by Lanczos; lambda returned too. */
int chs_fsai_omega(const chs_mat *A, const chs_mat *G, uint64_t seed, int nthreads,
                   double *omega, double *lambda);
/* Transpose and Galerkin product P^T (A P) (exact zeros dropped). */
int chs_transpose(const chs_mat *M, chs_mat **MT);
int chs_galerkin(const chs_mat *A, const chs_mat *P, int nthreads, chs_mat **Ac);
/* computeAMG.m:5-10 rule: diagonal entries equal to 1 become lambda_max(A)/10. */
int chs_bc_fix(const chs_mat *A, uint64_t seed, chs_mat **Afix, int32_t *nfixed);

/* ---- the full hierarchy ------------------------------------------------------------- */
/* Builds levels l = 0..nlev-1.  Level l < nlev-1 holds A_l, P_l, G_l, omega_l; the last
 * level holds only A_c (its factorisation is done by the consumer). */
int chs_build(const chs_mat *A, const chs_params *p, chs_hier **out);

/* Host replay of the GPU library's sliced-ELL single-vector product (no reference
 * counterpart): builds the layout the GPU library would upload for this CSR matrix and
 * walks it with the kernel's index arithmetic, checking every access against the layout's
 * buffers; y = A x with the kernel's summation order and rounding (fp32 != 0: FP32 values,
 * vector, products and sums).  info[6] = {code bits, sorted, log2 dictionary window, slices, log2 sort window, uniform
 * slice length or 0}; elems =
 * stored elements.  Returns 0, 1 when the matrix is not a sliced-ELL candidate, or a
 * negative code naming the first out-of-range access. */
int chs_sell_apply(int32_t n, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                   const double *val, int32_t fp32, const double *x, double *y, int32_t *info,
                   int64_t *elems);
/* Host replay of the GPU library's packed-code stream product (no reference counterpart):
 * builds the packed layout for PER nonzeros per thread and walks it with the kernel's
 * index arithmetic and summation order, checking every access; y = A x.  nblk receives the
 * block count.  Returns 0, 1 when the matrix is not a packed candidate, or a negative code
 * naming the first out-of-range access. */
int chs_pack_apply(int32_t n, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                   const double *val, int32_t per, int32_t fp32, const double *x, double *y,
                   int32_t *nblk);
int chs_hier_nlev(const chs_hier *h);
/* which: 'A', 'P' or 'G'.  Returned matrix is borrowed (owned by the hierarchy). */
const chs_mat *chs_hier_mat(const chs_hier *h, int32_t level, char which);
double chs_hier_omega(const chs_hier *h, int32_t level);
double chs_hier_setup_seconds(const chs_hier *h, int32_t phase);
void chs_hier_free(chs_hier *h);

#ifdef __cplusplus
}
#endif
#endif
