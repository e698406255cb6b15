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
 *   chs_soc_dom               sources/Preconditioner/AMG/coarsen/cpt_SoC_dom.m:1-80
 *   chs_agg_coarsen           sources/Preconditioner/AMG/coarsen/agg_coarsen.m:1-100,
 *                             mk_SoC_FcC.m:1-20
 *   chs_prol_bamg             sources/Preconditioner/AMG/prolong/MEX_BAMG_Prol/BAMG.cpp:16-241
 *                             (called by cpt_ProlMEX_BAMG.m:1-70 from prolong.m:37-41)
 *   chs_emin_pattern          sources/Preconditioner/AMG/prolong/EMIN/MEX_EMIN/
 *                             MEX_cptProlPatt.cpp:239-430 (prolong.m:213-228)
 *   chs_emin_improve          sources/Preconditioner/AMG/prolong/EMIN/Cpp_EMIN/
 *                             EMIN_matfree.cpp:61-300, gather_B_QR.cpp:30-480,
 *                             DEFL_PCG_matfree.cpp:28-250 (MEX_EMIN_enhance_prol.m:1-105,
 *                             prolong.m:231)
 *   chs_orthonormalize        sources/Preconditioner/AMG/tspace/tspace.m ('none': qr(V0,0)),
 *                             cpt_aspAMG.m:29-46
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
 *   chs_bc_fix                sources/utils/computeAMG.m:5-10,
 *                             sources/Drivers/driver_AMG/aspAMG_main.m:89-90
 *   chs_gen_rbm               sources/utils/mk_rbm_3d.m:1-27
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
  CHS_ECONSIST = -6, /* F/C count inconsistency (reference: cpt_aspAMG.m:212-215) */
  CHS_ENUM = -7      /* singular least-squares fit (reference: "error in LAPACKE_dgels") */
};

/* Opaque CSR matrix owned by the library. */
typedef struct chs_mat chs_mat;
/* Opaque hierarchy owned by the library. */
typedef struct chs_hier chs_hier;

/* Replacement of the aFSAI factor computation inside chs_build (e.g. the GPU kernel of
 * libchronos_amg, chr_setup_afsai in chronos_amg.h).  Contract of chr_afsai: A in CSR,
 * outputs rowptr_G[n+1] and col_G / val_G of n * (nstep*step_size + 1) entries, nnz_G the
 * entries written; returns 0, a negative code, or 1 when the level's rows are longer than
 * the replacement handles (chs_build then factors that level on the host). */
typedef int (*chs_afsai_fn)(void *ctx, int32_t n, const int32_t *rowptr, const int32_t *col,
                            const double *val, int32_t nstep, int32_t step_size, double eps,
                            int32_t *rowptr_G, int32_t *col_G, double *val_G, int64_t *nnz_G,
                            int32_t *is_posdef);

/* Storage callback for results sized at the end (see chs_galerkin_fn). */
typedef int (*chs_reserve_fn)(void *sink, int64_t nnz, int32_t **col, double **val);

/* Replacement of the Galerkin product inside chs_build (e.g. chr_setup_galerkin of
 * libchronos_amg): A_c = P^T (A P) for the level's A (n x n) and P (n x nc), written as
 * rowptr_C[nc+1] plus col/val storage obtained from reserve(sink, nnz, ...).  Returns 0,
 * a negative code, or 1 when it declines (chs_build then forms the product on the host). */
typedef int (*chs_galerkin_fn)(void *ctx, int32_t n, const int32_t *rowptr_A,
                               const int32_t *col_A, const double *val_A, int32_t nc,
                               const int32_t *rowptr_P, const int32_t *col_P,
                               const double *val_P, int32_t *rowptr_C, void *sink,
                               chs_reserve_fn reserve);

/* Replacement of EMIN's CG operator inside chs_build (e.g. chr_setup_emin_open / _apply /
 * _close of libchronos_amg): y = (A X) restricted to the pattern, X and y on the pattern
 * (trp, tci: the fine rows' pattern, n rows), each entry summed over A's row in order like
 * the host product.  open prepares a level (0, or nonzero to decline: the host then runs
 * the product), apply runs one product, close releases the level. */
typedef struct chs_emin_kp_ops {
  int (*open)(void *ctx, int32_t n, int32_t nc, const int32_t *rowptr_A, const int32_t *col_A,
              const double *val_A, const int32_t *rowptr_T, const int32_t *col_T, void **handle);
  int (*apply)(void *handle, const double *x, double *y);
  void (*close)(void *handle);
  void *ctx;
} chs_emin_kp_ops;

/* Replacement of the smoother damping computation inside chs_build (e.g. chr_setup_omega
 * of libchronos_amg): omega = min(1, 1.9 / lambda_max(G A G^T)) for the level's A and G. */
typedef int (*chs_omega_fn)(void *ctx, int32_t n, const int32_t *rowptr_A, const int32_t *col_A,
                            const double *val_A, const int32_t *rowptr_G, const int32_t *col_G,
                            const double *val_G, uint64_t seed, double *omega, double *lambda);

/* Setup parameters: the fields of the reference's parm_AMG / parm_COARSEN / parm_SMOOTH
 * files that the classical + aFSAI path reads (Tests/parm_*, examples/AMG/parm_*). */
/* parm_PROLONG's BAMG fields (Tests/parm_PROLONG lines 4-10; defaults in brackets).  The
 * number of interpolation points per row is bounded by the number of test vectors
 * (cpt_ProlMEX_BAMG.m:45). */
typedef struct chs_bamg_params {
  int32_t itmax_vol;    /* maximum volume iterations [100] */
  double tol_vol;       /* maximum volume tolerance [0.01] */
  int32_t dist_min;     /* belts of strong neighbours before the first fit [1] */
  int32_t dist_max;     /* belts at most [10] */
  double maxcond;       /* conditioning bound of the selected test-vector rows [1e13] */
  double maxrownrm;     /* bound on a row's weight 2-norm [5] */
  double eps_prol;      /* relative fit residual taken as exact [1e-10] */
} chs_bamg_params;

/* Energy minimisation of the prolongation (parm_PROLONG lines 12-17; prec_emin = 1 Jacobi
 * and solv_emin = 4 matrix product are the variants restated). */
typedef struct chs_emin_params {
  int32_t patt_pow;     /* powers of the strength pattern grown onto P's pattern [2] */
  double patt_tau;      /* DOM strength threshold of that pattern [0.01] */
  int32_t nnzr_max;     /* pattern entries per fine row at most, at least ntv [30] */
  int32_t itmax;        /* itmax_emin: CG iterations at most [10] */
  double energ_tol;     /* stop once an iteration's energy decrease falls below energ_tol
                           times the first one's [1e-3] */
  double condmax;       /* condmax_emin: conditioning bound of a row's constraint block [1e10] */
} chs_emin_params;

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
  int32_t prolong;      /* parm_PROLONG proltype: 0 = CLAS, 1 = EXTI, 2 = BAMG */
  int32_t nthreads;     /* OpenMP threads (<=0: runtime default) */
  uint64_t seed;        /* PMIS tie-break stream and Lanczos start vector */
  int32_t verbose;
  chs_afsai_fn afsai_fn; /* NULL: chs_afsai on the host */
  void *afsai_ctx;       /* passed to afsai_fn */
  chs_omega_fn omega_fn; /* NULL: chs_fsai_omega on the host */
  void *omega_ctx;       /* passed to omega_fn */
  int32_t soc_type;      /* parm_COARSEN SoC_type: 0 = CLA, 1 = DOM (tau in soc_tau) */
  int32_t ntv;           /* parm_TSPACE ntv: test vectors (BAMG) */
  int32_t ntv0;          /* columns of tv */
  const double *tv;      /* finest-level test space, n x ntv0 column-major; columns beyond
                            ntv0 (or all, when NULL) are uniform [0,1) draws of stream 7 */
  chs_bamg_params bamg;  /* prolong = 2 */
  int32_t prol_smooth;   /* parm_PROLONG prol_emin: 0 = NONE, 1 = SMOOTH (P - 1.9/lambda_max
                            D^-1 A P, prolong.m:180-197), 2 = EMIN (prolong.m:207-232, with
                            the level's test space; needs prolong = 2) */
  chs_galerkin_fn galerkin_fn; /* NULL: chs_galerkin on the host */
  void *galerkin_ctx;          /* passed to galerkin_fn */
  int32_t nl_agg;        /* parm_COARSEN nl_agg: levels 1..nl_agg use aggressive coarsening
                            (agg_coarsen.m) with the two-stage EXTI (or BAMG) prolongation */
  chs_emin_params emin;  /* prol_smooth = 2 */
  chs_emin_kp_ops emin_kp; /* open = NULL: EMIN's CG operator on the host */
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
/* Linear elasticity, trilinear hexahedra on a unit cube of ne^3 elements: (ne+1)^3 nodes x
 * 3 dofs (dof 3*node + component, node = ix + m*(iy + m*iz), m = ne+1), Young's modulus
 * E, Poisson ratio nu, 2x2x2 Gauss rule, all 27 neighbour nodes x 3 components in each
 * row.  rhs (may be NULL, n entries): uniform body force -1 in z.  clamp != 0: the z = 0
 * face's dofs become unit rows and columns with zero load. */
int chs_gen_elasticity(int32_t ne, double E, double nu, int32_t clamp, chs_mat **out,
                       double *rhs);
/* Rigid-body modes of those nodes about the cube's centroid (mk_rbm_3d.m): column-major
 * n x 6, translations x, y, z then rotations. */
void chs_gen_rbm(int32_t ne, double *rbm);

/* ---- setup components ------------------------------------------------------------- */
/* Classical strength of connection.  strong[e] = +1 strong / -1 weak for every stored
 * entry e of A (diagonal forced to +1), i.e. the sign pattern S_patt aligned with A. */
int chs_soc_cla(const chs_mat *A, double tau, int8_t *strong);
/* Strength by diagonal dominance: s_ij = |a_ij| / sqrt(a_ii a_jj), strong when s_ij > tau
 * (tau >= 0; the reference's tau < 0 variant, a kept-entry count, is not restated).  Same
 * output convention as chs_soc_cla. */
int chs_soc_dom(const chs_mat *A, double tau, int8_t *strong);
/* Aggressive coarsening (agg_coarsen.m): PMIS on the strength pattern (draws of stream
 * 100 + level), then PMIS over the coarse candidates' distance-two pattern |S|+|S^T| (draws
 * of stream 300 + level, no Dirichlet marking).  kind[i] = 2 coarse, 1 candidate left fine
 * (Fc), 0 fine (Ff).  chs_build uses it on levels 1..nl_agg with the two-stage EXTI
 * prolongation P = P1 P2 (cpt_TwoStageProlMEX_ExtI.m). */
int chs_agg_coarsen(const chs_mat *A, const int8_t *strong, uint64_t seed, int32_t level,
                    int8_t *kind);
/* Bootstrap-AMG least-squares prolongation.  Strength pattern: the entries of A with
 * strong[e] > 0.  tv: n x ntv column-major test space.  P (n x nc) interpolates every fine
 * node from at most ntv coarse nodes of its neighbourhood, grown belt by belt (dist_min to
 * dist_max), with the weights fitting the test vectors in least squares.  c_mark (n entries,
 * may be NULL) flags the fine rows whose fit stays inexact; *nmarked counts them. */
int chs_prol_bamg(const chs_mat *A, const int8_t *strong, const int32_t *fcnode, int32_t nc,
                  int32_t ntv, const double *tv, const chs_bamg_params *bp, int nthreads,
                  chs_mat **P, int32_t *c_mark, int32_t *nmarked);
/* Orthonormal basis of the columns of V (n x k column-major, in place), the Q of a thin QR
 * with positive R diagonal (tspace method NONE).  Returns 1 if the columns are dependent. */
int chs_orthonormalize(int64_t n, int32_t k, double *V);
/* EMIN prolongation pattern (MEX_cptProlPatt.cpp:239-430, called at prolong.m:228): the
 * fine rows of P's pattern (stored zeros ignored) grown to at most max(nnzr_max, ntv)
 * entries by the largest |(S^(k-1) P)_ij| outside it, ties to the lower column, S the DOM
 * strength matrix of A above patt_tau with values |a_ij|/sqrt(a_ii a_jj); coarse rows are
 * empty.  *patt: n x nc, values 1. */
int chs_emin_pattern(const chs_mat *A, const chs_mat *P, const int32_t *fcnode, int32_t ntv,
                     const chs_emin_params *prm, int nthreads, chs_mat **patt);
/* EMIN (MEX_EMIN_enhance_prol.m -> EMIN_Prolong_compute.cpp -> EMIN_matfree.cpp:61-300 with
 * solv_emin = 4, prec_emin = 1; symmetric A): P's fine rows copied into the pattern, each
 * row corrected onto its test-space constraint P_i TV_C = TV_i (QR, SVD below full rank;
 * gather_B_QR.cpp), then the projected Jacobi-PCG on trace(P^T A P) over the pattern
 * (DEFL_PCG_matfree.cpp) for at most itmax iterations.  tv: n x ntv column-major.  *Pout:
 * n x nc with coarse rows e_c and exact zeros dropped; *iter: CG iterations. */
int chs_emin_improve(const chs_mat *A, const chs_mat *P, const chs_mat *patt,
                     const int32_t *fcnode, int32_t ntv, const double *tv,
                     const chs_emin_params *prm, int nthreads, chs_mat **Pout, int32_t *iter);
/* chs_emin_improve with the CG operator from ops (see chs_emin_kp_ops; NULL or a declined
 * open: the host product).  The same result bit for bit. */
int chs_emin_improve_ops(const chs_mat *A, const chs_mat *P, const chs_mat *patt,
                         const int32_t *fcnode, int32_t ntv, const double *tv,
                         const chs_emin_params *prm, const chs_emin_kp_ops *ops, int nthreads,
                         chs_mat **Pout, int32_t *iter);

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
/* The factor repair of sources/Preconditioner/AMG/smoother/smoother.m:59-71: every row of G
 * whose diagonal entry is NaN becomes the single diagonal entry 1/sqrt(a_ii) (in place);
 * nfixed = rows repaired.  chs_build applies it to every aFSAI factor. */
int chs_fsai_nan_repair(const chs_mat *A, chs_mat *G, int32_t *nfixed);
/* omega = min(1, 1.9 / lambda_max(G A G^T)) by Lanczos; lambda returned too. */
int chs_fsai_omega(const chs_mat *A, const chs_mat *G, uint64_t seed, int nthreads,
                   double *omega, double *lambda);
/* Transpose and Galerkin product P^T (A P) (exact zeros dropped). */
int chs_transpose(const chs_mat *M, chs_mat **MT);
int chs_galerkin(const chs_mat *A, const chs_mat *P, int nthreads, chs_mat **Ac);
/* Dirichlet-row rescaling with lambda_max(A) (Lanczos, start vector from seed):
 * CHS_BC_DIAG: diagonal entries equal to 1 become lambda_max/10 (sources/utils/computeAMG.m:5-10);
 * CHS_BC_ALL: every stored entry equal to 1 becomes lambda_max/10
 * (sources/Drivers/driver_AMG/aspAMG_main.m:89-90).  nfixed = entries changed.  nthreads <= 0:
 * the OpenMP default; the result does not depend on it. */
enum { CHS_BC_DIAG = 0, CHS_BC_ALL = 1 };
int chs_bc_fix(const chs_mat *A, uint64_t seed, int32_t mode, int nthreads, chs_mat **Afix,
               int32_t *nfixed);

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
/* Host replay of the GPU library's symmetric diagonal product (no reference counterpart):
 * builds the layout the GPU library would upload (the main and upper diagonals of a matrix
 * with at most 16 of them, values symmetric bit for bit, no stored zeros, sorted rows) and
 * walks it with the kernel's index arithmetic and summation order; y = A x (fp32 != 0:
 * FP32 values, vector, products and sums).  ndiag receives the stored diagonal count and
 * offsets (room for 16) their offsets.  Returns 0, 1 when the matrix is not a candidate,
 * or a negative code naming the first out-of-range access. */
int chs_dia_apply(int32_t n, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                  const double *val, int32_t fp32, const double *x, double *y, int32_t *ndiag,
                  int32_t *offsets);
/* The host dense routines of the GPU library's LOBPCG Rayleigh-Ritz steps, exported for
 * testing.  chs_sygv: A v = lambda B v for symmetric A and SPD B (order m, column-major),
 * ascending eigenvalues w and B-orthonormal eigenvectors V (MATLAB's [V, L] = eig(A, B) in
 * sources/EigenSolver/lobpcg.m:99, :325, :437); returns 0, 1 when B is not positive definite,
 * -1 on bad arguments.  chs_cond_sym: the 2-norm condition number of a symmetric matrix
 * (cond in sources/EigenSolver/B_orthonormalize.m:9); -1 on bad arguments. */
int chs_sygv(int32_t m, const double *A, const double *B, double *w, double *V);
double chs_cond_sym(int32_t m, const double *M);
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
