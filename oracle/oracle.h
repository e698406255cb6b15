/*
 * oracle.h — CPU FP64 restatement of ChronosLab's AMG solve phase.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP solve phase is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product never links or calls it.
 *
 * Restated functions (reference paths relative to the reference root):
 *   orc_vcycle     sources/Preconditioner/AMG/aspAMG/AMG_Vcycle.m:1-54 with the smoother
 *                  closures of cpt_aspAMG.m:301-302 and the coarsest solve of :103-131
 *   orc_pcg        sources/LinearSolver/PCG_CJ.m:1-73
 *   orc_lobpcg     sources/EigenSolver/lobpcg.m:21-467 + B_orthonormalize.m:1-25
 *                  (prodB = [] and an optional constraint block Y, as
 *                  sources/Drivers/driver_Eigen/EigDriver.m:202-226 calls it)
 *   orc_fv_grad    no reference counterpart: gradient of 0.5||x - x_obs||^2 with respect to
 *                  ln k for the finite-volume diffusion operator, from the adjoint field.
 * All dense vectors are column-major n x k blocks (MATLAB layout).
 */
#ifndef CHRONOS_ORACLE_H
#define CHRONOS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_csr {
  int32_t nrows, ncols;
  const int32_t *rowptr, *col;
  const double *val;
} orc_csr;

typedef struct orc_level {
  orc_csr A, P, G;
  double omega;
  int32_t nupre, nupost;
} orc_level;

typedef struct orc_hier orc_hier;

int orc_hier_create(int32_t nlev, const orc_level *levels, orc_hier **out);
void orc_hier_free(orc_hier *h);
void orc_set_threads(int32_t nthreads);

int orc_spmv(const orc_csr *A, const double *x, double *y);
int orc_vcycle(const orc_hier *h, int32_t k, const double *b, int64_t ldb, double *x,
               int64_t ldx);
/* prodA = A (the outer operator), prodM = V-cycle of h.  x holds x0 on entry. */
int orc_pcg(const orc_hier *h, const orc_csr *A, const double *b, double *x, int32_t itmax,
            double tol, int32_t *iter, double *resvec, int32_t *flag);
/* prodM = V-cycle of h, or the identity when h is NULL.  X (n x neig) holds X0 on entry
 * and the eigenvectors on exit; lambda[neig]; histories may be NULL. */
/* Y: the constraint block, n x m column-major (m = 0: none); returns -2 when Y^T Y is not
 * positive definite. */
int orc_lobpcg(const orc_csr *A, const orc_hier *h, double *X, const double *Y, int32_t m,
               int32_t n, int32_t neig, int32_t largest, int32_t reslambda_check, int32_t itmax,
               double tol, int32_t restart_control, int32_t *niter, double *lambda,
               double *resnorm_hist, double *lambda_hist, int32_t *ierr);
/* dL/d(ln k) for the 7-point finite-volume operator: A_ij = -2 k_i k_j/(k_i+k_j), plus
 * Dirichlet terms 2 k_i on the x-faces, with the adjoint lam solving A lam = x - x_obs. */
int orc_fv_grad(int32_t N, const double *k, const double *x, const double *lam, double *g);
/* Dense symmetric-definite generalized eigensolver (Cholesky reduction + cyclic Jacobi),
 * ascending eigenvalues, B-orthonormal eigenvectors; exported for tests. */
int orc_sygv(int32_t m, const double *A, const double *B, double *w, double *V);

#ifdef __cplusplus
}
#endif
#endif
