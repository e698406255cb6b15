/*
 * CPU FP64 restatement of ChronosLab's AMG solve phase — TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle that the HIP solve phase is checked against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; no product code
 * links or calls it.  See oracle.h for the restated reference functions.
 *
 * Parity pins (tests/test_oracle.py): V-cycle linearity and symmetry, the dense two-level
 * closed form of the V(1,1) operator, the exact single-level solve, agreement of orc_pcg
 * with a textbook CG (scipy) for a fixed SPD preconditioner, LOBPCG eigenvalues against the
 * closed-form Dirichlet Laplacian spectrum and numpy.linalg.eigh, orc_sygv against
 * scipy.linalg.eigh, and orc_fv_grad against central finite differences.
 *
 * Dot products are summed over a fixed partition of 64 chunks and the partials are added in
 * chunk order, so results do not depend on the OpenMP thread count.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#define NCHUNK 64

static int g_threads = 0;
void orc_set_threads(int32_t nthreads) { g_threads = nthreads; }
static int nth(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* ---------------------------------------------------------------------------------- */
/* Vector kernels                                                                      */
/* ---------------------------------------------------------------------------------- */
static double dot(int64_t n, const double *x, const double *y) {
  double part[NCHUNK];
#pragma omp parallel for num_threads(nth()) schedule(static)
  for (int c = 0; c < NCHUNK; ++c) {
    int64_t lo = n * c / NCHUNK, hi = n * (c + 1) / NCHUNK;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s += x[i] * y[i];
    part[c] = s;
  }
  double s = 0.0;
  for (int c = 0; c < NCHUNK; ++c) s += part[c];
  return s;
}

static double nrm2(int64_t n, const double *x) { return sqrt(dot(n, x, x)); }

static void spmv(const orc_csr *A, const double *x, double *y) {
#pragma omp parallel for num_threads(nth()) schedule(static)
  for (int32_t i = 0; i < A->nrows; ++i) {
    double s = 0.0;
    for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e) s += A->val[e] * x[A->col[e]];
    y[i] = s;
  }
}

int orc_spmv(const orc_csr *A, const double *x, double *y) {
  if (!A || !x || !y) return -1;
  spmv(A, x, y);
  return 0;
}

/* ---------------------------------------------------------------------------------- */
/* Hierarchy                                                                           */
/* ---------------------------------------------------------------------------------- */
typedef struct {
  int32_t nrows, ncols;
  int32_t *rowptr, *col;
  double *val;
} owned_csr;

typedef struct {
  orc_csr A, P, G;
  owned_csr PT, GT;
  double omega;
  int32_t nupre, nupost;
  double *r, *t, *u, *bc, *xc;
} olevel;

struct orc_hier {
  int32_t nlev;
  olevel *lev;
  int32_t nc;
  double *L; /* dense lower Cholesky factor of the coarsest operator, column-major */
};

static void transpose(const orc_csr *M, owned_csr *T) {
  T->nrows = M->ncols;
  T->ncols = M->nrows;
  T->rowptr = (int32_t *)calloc((size_t)T->nrows + 1, sizeof(int32_t));
  int64_t nnz = M->rowptr[M->nrows];
  T->col = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  T->val = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
  for (int64_t e = 0; e < nnz; ++e) T->rowptr[M->col[e] + 1]++;
  for (int32_t i = 0; i < T->nrows; ++i) T->rowptr[i + 1] += T->rowptr[i];
  int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * ((size_t)T->nrows + 1));
  memcpy(pos, T->rowptr, sizeof(int32_t) * ((size_t)T->nrows + 1));
  for (int32_t i = 0; i < M->nrows; ++i)
    for (int32_t e = M->rowptr[i]; e < M->rowptr[i + 1]; ++e) {
      int32_t p = pos[M->col[e]]++;
      T->col[p] = i;
      T->val[p] = M->val[e];
    }
  free(pos);
}

static orc_csr view(const owned_csr *m) {
  orc_csr v = {m->nrows, m->ncols, m->rowptr, m->col, m->val};
  return v;
}

static void free_owned(owned_csr *m) {
  free(m->rowptr);
  free(m->col);
  free(m->val);
}

/* Dense lower Cholesky, column-major, in place.  Returns 0 or the failing column + 1. */
static int chol_dense(int32_t m, double *K) {
  for (int32_t j = 0; j < m; ++j) {
    double d = K[j + (int64_t)j * m];
    for (int32_t k = 0; k < j; ++k) d -= K[j + (int64_t)k * m] * K[j + (int64_t)k * m];
    if (!(d > 0.0)) return j + 1;
    d = sqrt(d);
    K[j + (int64_t)j * m] = d;
    for (int32_t i = j + 1; i < m; ++i) {
      double s = K[i + (int64_t)j * m];
      for (int32_t k = 0; k < j; ++k) s -= K[i + (int64_t)k * m] * K[j + (int64_t)k * m];
      K[i + (int64_t)j * m] = s / d;
    }
    for (int32_t i = 0; i < j; ++i) K[i + (int64_t)j * m] = 0.0;
  }
  return 0;
}

static void chol_dense_solve(int32_t m, const double *L, double *b) {
  for (int32_t i = 0; i < m; ++i) {
    double s = b[i];
    for (int32_t k = 0; k < i; ++k) s -= L[i + (int64_t)k * m] * b[k];
    b[i] = s / L[i + (int64_t)i * m];
  }
  for (int32_t i = m - 1; i >= 0; --i) {
    double s = b[i];
    for (int32_t k = i + 1; k < m; ++k) s -= L[k + (int64_t)i * m] * b[k];
    b[i] = s / L[i + (int64_t)i * m];
  }
}

int orc_hier_create(int32_t nlev, const orc_level *levels, orc_hier **out) {
  if (nlev < 1 || !levels || !out) return -1;
  orc_hier *h = (orc_hier *)calloc(1, sizeof(orc_hier));
  h->nlev = nlev;
  h->lev = (olevel *)calloc((size_t)nlev, sizeof(olevel));
  for (int32_t l = 0; l < nlev; ++l) {
    olevel *L = &h->lev[l];
    L->A = levels[l].A;
    L->omega = levels[l].omega;
    L->nupre = levels[l].nupre;
    L->nupost = levels[l].nupost;
    int32_t n = L->A.nrows;
    if (l + 1 < nlev) {
      L->P = levels[l].P;
      L->G = levels[l].G;
      transpose(&L->P, &L->PT);
      transpose(&L->G, &L->GT);
      int32_t nc = L->P.ncols;
      L->r = (double *)malloc(sizeof(double) * n);
      L->t = (double *)malloc(sizeof(double) * n);
      L->u = (double *)malloc(sizeof(double) * n);
      L->bc = (double *)malloc(sizeof(double) * nc);
      L->xc = (double *)malloc(sizeof(double) * nc);
    }
  }
  /* Coarsest level: dense Cholesky (cpt_aspAMG.m:109 factors the sparse matrix with a
   * fill-reducing permutation; the permutation does not change the solution). */
  const orc_csr *Ac = &h->lev[nlev - 1].A;
  int32_t m = Ac->nrows;
  if ((int64_t)m * m > (int64_t)1 << 27) {
    orc_hier_free(h);
    return -2;
  }
  h->nc = m;
  h->L = (double *)calloc((size_t)m * m, sizeof(double));
  for (int32_t i = 0; i < m; ++i)
    for (int32_t e = Ac->rowptr[i]; e < Ac->rowptr[i + 1]; ++e)
      h->L[i + (int64_t)Ac->col[e] * m] = Ac->val[e];
  if (chol_dense(m, h->L)) {
    orc_hier_free(h);
    return -4;
  }
  *out = h;
  return 0;
}

void orc_hier_free(orc_hier *h) {
  if (!h) return;
  for (int32_t l = 0; l < h->nlev; ++l) {
    olevel *L = &h->lev[l];
    if (l + 1 < h->nlev) {
      free_owned(&L->PT);
      free_owned(&L->GT);
      free(L->r);
      free(L->t);
      free(L->u);
      free(L->bc);
      free(L->xc);
    }
  }
  free(h->lev);
  free(h->L);
  free(h);
}

/* z = omega * G^T (G r)  (cpt_aspAMG.m:301-302) */
static void smooth(olevel *L, const double *r, double *z) {
  orc_csr GT = view(&L->GT);
  spmv(&L->G, r, L->t);
  spmv(&GT, L->t, z);
  const int32_t n = L->A.nrows;
  for (int32_t i = 0; i < n; ++i) z[i] *= L->omega;
}

/* r = b - A x */
static void residual(const orc_csr *A, const double *b, const double *x, double *r) {
#pragma omp parallel for num_threads(nth()) schedule(static)
  for (int32_t i = 0; i < A->nrows; ++i) {
    double s = 0.0;
    for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e) s += A->val[e] * x[A->col[e]];
    r[i] = b[i] - s;
  }
}

/* AMG_Vcycle.m:1-54 with a zero initial guess (nargin < 4 at every level). */
static void vcycle_rec(const orc_hier *h, int32_t l, const double *b, double *x) {
  olevel *L = &h->lev[l];
  const int32_t n = L->A.nrows;
  if (l == h->nlev - 1) {
    memcpy(x, b, sizeof(double) * n);
    chol_dense_solve(h->nc, h->L, x);
    return;
  }
  smooth(L, b, x);
  for (int32_t k = 2; k <= L->nupre; ++k) {
    residual(&L->A, b, x, L->r);
    smooth(L, L->r, L->u);
    for (int32_t i = 0; i < n; ++i) x[i] += L->u[i];
  }
  residual(&L->A, b, x, L->r);
  orc_csr PT = view(&L->PT);
  spmv(&PT, L->r, L->bc);
  vcycle_rec(h, l + 1, L->bc, L->xc);
  spmv(&L->P, L->xc, L->u);
  for (int32_t i = 0; i < n; ++i) x[i] += L->u[i];
  for (int32_t k = 1; k <= L->nupost; ++k) {
    residual(&L->A, b, x, L->r);
    smooth(L, L->r, L->u);
    for (int32_t i = 0; i < n; ++i) x[i] += L->u[i];
  }
}

int orc_vcycle(const orc_hier *h, int32_t k, const double *b, int64_t ldb, double *x,
               int64_t ldx) {
  if (!h || !b || !x || k < 0) return -1;
  for (int32_t c = 0; c < k; ++c) vcycle_rec(h, 0, b + c * ldb, x + c * ldx);
  return 0;
}

/* ---------------------------------------------------------------------------------- */
/* PCG_CJ.m:1-73                                                                       */
/* ---------------------------------------------------------------------------------- */
int orc_pcg(const orc_hier *h, const orc_csr *A, const double *b, double *x, int32_t itmax,
            double tol, int32_t *iter_out, double *resvec, int32_t *flag_out) {
  if (!h || !A || !b || !x) return -1;
  const int32_t n = A->nrows;
  int32_t iter = 0, flag = 0;
  double bnorm = nrm2(n, b);
  if (bnorm == 0.0) {
    memset(x, 0, sizeof(double) * n);
    if (resvec) resvec[0] = 0.0;
    if (iter_out) *iter_out = 0;
    if (flag_out) *flag_out = 0;
    return 0;
  }
  double *res = (double *)malloc(sizeof(double) * n);
  double *pres = (double *)malloc(sizeof(double) * n);
  double *p = (double *)calloc(n, sizeof(double));
  double *axp = (double *)calloc(n, sizeof(double));
  residual(A, b, x, res);
  double resiter = nrm2(n, res) / bnorm;
  if (resvec) resvec[0] = resiter;
  int exit_test = itmax <= 0;
  double ptap = 0.0;
  while (!exit_test) {
    ++iter;
    vcycle_rec(h, 0, res, pres);
    if (iter == 1) {
      memcpy(p, pres, sizeof(double) * n);
    } else {
      double beta = -dot(n, pres, axp) / ptap;
      for (int32_t i = 0; i < n; ++i) p[i] = pres[i] + beta * p[i];
    }
    spmv(A, p, axp);
    ptap = dot(n, p, axp);
    double alpha = dot(n, p, res) / ptap;
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      res[i] -= alpha * axp[i];
    }
    resiter = nrm2(n, res) / bnorm;
    if (resvec) resvec[iter] = resiter;
    exit_test = resiter < tol || iter == itmax;
  }
  if (resiter > tol) flag = 1;
  free(res);
  free(pres);
  free(p);
  free(axp);
  if (iter_out) *iter_out = iter;
  if (flag_out) *flag_out = flag;
  return 0;
}

/* ---------------------------------------------------------------------------------- */
/* Dense helpers for LOBPCG (column-major)                                             */
/* ---------------------------------------------------------------------------------- */
/* C (a x b) = X^T Y with X n x a, Y n x b; chunked deterministic sums. */
static void gemm_tn(int64_t n, int32_t a, const double *X, int32_t b, const double *Y,
                    double *Cm) {
  for (int32_t j = 0; j < b; ++j)
    for (int32_t i = 0; i < a; ++i) Cm[i + (int64_t)j * a] = dot(n, X + i * n, Y + j * n);
}

/* Z (n x b) = X (n x a) * Cm (a x b) */
static void gemm_nn(int64_t n, int32_t a, const double *X, int32_t b, const double *Cm,
                    double *Z) {
#pragma omp parallel for num_threads(nth()) schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int32_t j = 0; j < b; ++j) {
      double s = 0.0;
      for (int32_t k = 0; k < a; ++k) s += X[i + k * n] * Cm[k + (int64_t)j * a];
      Z[i + j * n] = s;
    }
}

static void symmetrize(int32_t m, double *M) {
  for (int32_t j = 0; j < m; ++j)
    for (int32_t i = j + 1; i < m; ++i) {
      double s = 0.5 * (M[i + j * m] + M[j + i * m]);
      M[i + j * m] = s;
      M[j + i * m] = s;
    }
}

/* Cyclic Jacobi eigenvalue algorithm for a symmetric m x m matrix (Golub & Van Loan,
 * Alg. 8.5.3).  On exit w holds the eigenvalues (unsorted) and Q the eigenvectors. */
static void jacobi_eig(int32_t m, double *S, double *w, double *Q) {
  for (int32_t j = 0; j < m; ++j)
    for (int32_t i = 0; i < m; ++i) Q[i + j * m] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int32_t j = 0; j < m; ++j)
      for (int32_t i = 0; i < m; ++i) {
        double v = S[i + j * m] * S[i + j * m];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int32_t p = 0; p < m - 1; ++p)
      for (int32_t q = p + 1; q < m; ++q) {
        double apq = S[p + q * m];
        if (apq == 0.0) continue;
        double app = S[p + p * m], aqq = S[q + q * m];
        double tau = (aqq - app) / (2.0 * apq);
        double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int32_t k = 0; k < m; ++k) {
          double skp = S[k + p * m], skq = S[k + q * m];
          S[k + p * m] = c * skp - s * skq;
          S[k + q * m] = s * skp + c * skq;
        }
        for (int32_t k = 0; k < m; ++k) {
          double spk = S[p + k * m], sqk = S[q + k * m];
          S[p + k * m] = c * spk - s * sqk;
          S[q + k * m] = s * spk + c * sqk;
        }
        for (int32_t k = 0; k < m; ++k) {
          double qkp = Q[k + p * m], qkq = Q[k + q * m];
          Q[k + p * m] = c * qkp - s * qkq;
          Q[k + q * m] = s * qkp + c * qkq;
        }
      }
  }
  for (int32_t i = 0; i < m; ++i) w[i] = S[i + i * m];
}

int orc_sygv(int32_t m, const double *A, const double *B, double *w, double *V) {
  double *L = (double *)malloc(sizeof(double) * m * m);
  double *Cm = (double *)malloc(sizeof(double) * m * m);
  double *Q = (double *)malloc(sizeof(double) * m * m);
  double *wt = (double *)malloc(sizeof(double) * m);
  int32_t *ord = (int32_t *)malloc(sizeof(int32_t) * m);
  memcpy(L, B, sizeof(double) * m * m);
  int rc = chol_dense(m, L);
  if (rc) {
    free(L); free(Cm); free(Q); free(wt); free(ord);
    return -4;
  }
  /* Cm = L^-1 A L^-T : solve column by column. */
  memcpy(Cm, A, sizeof(double) * m * m);
  for (int32_t j = 0; j < m; ++j) { /* Cm(:,j) = L^-1 A(:,j) */
    double *c = Cm + j * m;
    for (int32_t i = 0; i < m; ++i) {
      double s = c[i];
      for (int32_t k = 0; k < i; ++k) s -= L[i + k * m] * c[k];
      c[i] = s / L[i + i * m];
    }
  }
  for (int32_t i = 0; i < m; ++i) { /* rows: Cm(i,:) = Cm(i,:) L^-T */
    for (int32_t j = 0; j < m; ++j) {
      double s = Cm[i + j * m];
      for (int32_t k = 0; k < j; ++k) s -= Cm[i + k * m] * L[j + k * m];
      Cm[i + j * m] = s / L[j + j * m];
    }
  }
  symmetrize(m, Cm);
  jacobi_eig(m, Cm, wt, Q);
  for (int32_t i = 0; i < m; ++i) ord[i] = i;
  for (int32_t i = 1; i < m; ++i) { /* insertion sort, ascending */
    int32_t t = ord[i], j = i - 1;
    while (j >= 0 && wt[ord[j]] > wt[t]) { ord[j + 1] = ord[j]; --j; }
    ord[j + 1] = t;
  }
  for (int32_t jj = 0; jj < m; ++jj) {
    int32_t j = ord[jj];
    w[jj] = wt[j];
    /* V(:,jj) = L^-T Q(:,j) */
    double *v = V + jj * m;
    for (int32_t i = 0; i < m; ++i) v[i] = Q[i + j * m];
    for (int32_t i = m - 1; i >= 0; --i) {
      double s = v[i];
      for (int32_t k = i + 1; k < m; ++k) s -= L[k + i * m] * v[k];
      v[i] = s / L[i + i * m];
    }
  }
  free(L); free(Cm); free(Q); free(wt); free(ord);
  return 0;
}

/* Condition number of a symmetric positive semi-definite matrix (|lambda| ratio). */
static double cond_sym(int32_t m, const double *M) {
  double *S = (double *)malloc(sizeof(double) * m * m);
  double *Q = (double *)malloc(sizeof(double) * m * m);
  double *w = (double *)malloc(sizeof(double) * m);
  memcpy(S, M, sizeof(double) * m * m);
  jacobi_eig(m, S, w, Q);
  double mx = 0.0, mn = DBL_MAX;
  for (int32_t i = 0; i < m; ++i) {
    mx = fmax(mx, fabs(w[i]));
    mn = fmin(mn, fabs(w[i]));
  }
  free(S); free(Q); free(w);
  return mn == 0.0 ? INFINITY : mx / mn;
}

/* B_orthonormalize.m:1-25 with prodB = []: X <- X U^-1, XBX = U^T U.  Returns 0 or -1
 * when cond(X^T X) exceeds 1/(k*eps).  U (k x k, column-major upper) is returned. */
static int b_orthonormalize(int64_t n, int32_t k, double *X, double *U) {
  double *M = (double *)malloc(sizeof(double) * k * k);
  gemm_tn(n, k, X, k, X, M);
  /* XBX = 0.5*(XBX+XBX)' in the reference: the transpose of XBX. */
  for (int32_t j = 0; j < k; ++j)
    for (int32_t i = 0; i < j; ++i) {
      double t = M[i + j * k];
      M[i + j * k] = M[j + i * k];
      M[j + i * k] = t;
    }
  if (cond_sym(k, M) > 1.0 / ((double)k * DBL_EPSILON)) {
    free(M);
    return -1;
  }
  /* chol upper: M = U^T U from the lower factor of M (uses the lower triangle). */
  if (chol_dense(k, M)) {
    free(M);
    return -1;
  }
  for (int32_t j = 0; j < k; ++j)
    for (int32_t i = 0; i < k; ++i) U[i + j * k] = (i <= j) ? M[j + i * k] : 0.0;
  free(M);
  /* X <- X U^-1, row by row (forward substitution on U^T). */
#pragma omp parallel for num_threads(nth()) schedule(static)
  for (int64_t r = 0; r < n; ++r)
    for (int32_t j = 0; j < k; ++j) {
      double s = X[r + j * n];
      for (int32_t i = 0; i < j; ++i) s -= X[r + i * n] * U[i + j * k];
      X[r + j * n] = s / U[j + j * k];
    }
  return 0;
}

/* Y <- Y U^-1 */
static void right_solve_upper(int64_t n, int32_t k, double *Y, const double *U) {
#pragma omp parallel for num_threads(nth()) schedule(static)
  for (int64_t r = 0; r < n; ++r)
    for (int32_t j = 0; j < k; ++j) {
      double s = Y[r + j * n];
      for (int32_t i = 0; i < j; ++i) s -= Y[r + i * n] * U[i + j * k];
      Y[r + j * n] = s / U[j + j * k];
    }
}

static void apply_A_block(const orc_csr *A, int64_t n, int32_t k, const double *X, double *Y) {
  for (int32_t j = 0; j < k; ++j) spmv(A, X + j * n, Y + j * n);
}

static void apply_M_block(const orc_hier *h, int64_t n, int32_t k, double *X, double *tmp) {
  if (!h) return;
  for (int32_t j = 0; j < k; ++j) {
    vcycle_rec(h, 0, X + j * n, tmp);
    memcpy(X + j * n, tmp, sizeof(double) * n);
  }
}

/* The constraint projector of lobpcg.m:64-75 with prodB = []: with U^T U = Y^T Y (chol),
 * X <- X - Y U^-1 U^-T (Y^T X).  Y: n x m column-major, Uy: m x m column-major upper. */
static void orthog_y(int64_t n, int32_t m, const double *Y, const double *Uy, int32_t k, double *X,
                     double *tmp) {
  double *C = (double *)malloc(sizeof(double) * m * k);
  gemm_tn(n, m, Y, k, X, C);
  for (int32_t j = 0; j < k; ++j) {
    double *c = C + (int64_t)j * m;
    for (int32_t i = 0; i < m; ++i) { /* U^T z = c */
      double s = c[i];
      for (int32_t q = 0; q < i; ++q) s -= Uy[q + i * m] * c[q];
      c[i] = s / Uy[i + i * m];
    }
    for (int32_t i = m - 1; i >= 0; --i) { /* U w = z */
      double s = c[i];
      for (int32_t q = i + 1; q < m; ++q) s -= Uy[i + q * m] * c[q];
      c[i] = s / Uy[i + i * m];
    }
  }
  gemm_nn(n, m, Y, k, C, tmp);
  for (int64_t i = 0; i < n * k; ++i) X[i] -= tmp[i];
  free(C);
}

/* lobpcg.m:21-467 with prodB = []; Y (n x m column-major, m = 0: none) is the constraint
 * block: X0, the preconditioned residuals and the final X are projected B-orthogonally to it
 * (lobpcg.m:64-75, 228-230, 384-387). */
int orc_lobpcg(const orc_csr *A, const orc_hier *h, double *X, const double *Y, int32_t m_y,
               int32_t n_, int32_t neig, int32_t largest, int32_t reslambda_check, int32_t itmax,
               double tol, int32_t restart_control, int32_t *niter_out, double *lambda_out,
               double *resnorm_hist, double *lambda_hist, int32_t *ierr_out) {
  const int64_t n = n_;
  if (!A || !X || neig < 1 || itmax < 2 || n < 3 * (int64_t)neig || m_y < 0 || (m_y > 0 && !Y)) return -1;
  double *Uy = NULL;
  if (m_y > 0) { /* YBY = Y^T Y; U_YBY = chol(YBY) (upper) */
    Uy = (double *)malloc(sizeof(double) * m_y * m_y);
    double *M = (double *)malloc(sizeof(double) * m_y * m_y);
    gemm_tn(n, m_y, Y, m_y, Y, M);
    if (chol_dense(m_y, M)) {
      free(M);
      free(Uy);
      return -2;
    }
    for (int32_t j = 0; j < m_y; ++j)
      for (int32_t i = 0; i < m_y; ++i) Uy[i + j * m_y] = (i <= j) ? M[j + i * m_y] : 0.0;
    free(M);
  }
  const size_t blk = sizeof(double) * n * neig;
  double *AX = (double *)malloc(blk), *P = (double *)calloc(n * neig, sizeof(double));
  double *AP = (double *)calloc(n * neig, sizeof(double)), *R = (double *)malloc(blk);
  double *AR = (double *)malloc(blk), *AAR = (double *)malloc(blk);
  double *APa = (double *)malloc(blk), *AAP = (double *)malloc(blk);
  double *best = (double *)malloc(blk), *tmp = (double *)malloc(blk);
  double *vtmp = (double *)malloc(sizeof(double) * n);
  const int32_t mmax = 3 * neig;
  double *gA = (double *)malloc(sizeof(double) * mmax * mmax);
  double *gB = (double *)malloc(sizeof(double) * mmax * mmax);
  double *E = (double *)malloc(sizeof(double) * mmax * mmax);
  double *wall = (double *)malloc(sizeof(double) * mmax);
  double *Csm = (double *)malloc(sizeof(double) * mmax * mmax);
  double *U = (double *)malloc(sizeof(double) * neig * neig);
  double lam[64], lam_old[64], relDelta[64], resn[64];
  int32_t act[64];
  int32_t ierr = 0, iter = 0, n_active = neig;
  if (neig > 64) return -1;
  for (int32_t j = 0; j < neig; ++j) {
    lam_old[j] = 0.0;
    relDelta[j] = INFINITY;
  }
#define HIST(hist, it, j) if (hist) (hist)[(int64_t)(j) * (itmax + 1) + (it)]
  double smallest = DBL_MAX;
  memcpy(best, X, blk);
  if (m_y > 0) orthog_y(n, m_y, Y, Uy, neig, X, tmp);
  if (b_orthonormalize(n, neig, X, U)) {
    ierr = 1;
    goto done;
  }
  apply_A_block(A, n, neig, X, AX);
  gemm_tn(n, neig, X, neig, AX, gA);
  symmetrize(neig, gA);
  gemm_tn(n, neig, X, neig, X, gB);
  symmetrize(neig, gB);
  if (orc_sygv(neig, gA, gB, wall, E)) { ierr = 1; goto done; }
  for (int32_t j = 0; j < neig; ++j) { lam[j] = wall[j]; HIST(lambda_hist, 0, j) = lam[j]; }
  gemm_nn(n, neig, X, neig, E, tmp); memcpy(X, tmp, blk);
  gemm_nn(n, neig, AX, neig, E, tmp); memcpy(AX, tmp, blk);
  int restart = 1, forced = 0;
  while (iter < itmax - 1) {
    ++iter;
    for (int32_t j = 0; j < neig; ++j)
      for (int64_t i = 0; i < n; ++i) R[i + j * n] = AX[i + j * n] - X[i + j * n] * lam[j];
    double avg = 0.0;
    for (int32_t j = 0; j < neig; ++j) {
      resn[j] = nrm2(n, R + j * n) / lam[j];
      HIST(resnorm_hist, iter - 1, j) = resn[j];
      avg += resn[j];
      relDelta[j] = fabs(lam_old[j] - lam[j]) / lam[j];
      lam_old[j] = lam[j];
    }
    avg /= neig;
    if (avg < smallest) {
      smallest = avg;
      memcpy(best, X, blk);
    } else if (avg > ldexp(1.0, restart_control) * smallest) {
      forced = 1;
      apply_A_block(A, n, neig, X, AX);
    }
    n_active = 0;
    for (int32_t j = 0; j < neig; ++j) {
      int conv = reslambda_check ? (relDelta[j] < tol) : (resn[j] < tol);
      if (!conv) act[n_active++] = j;
    }
    if (n_active == 0) break;
    const int32_t na = n_active;
    for (int32_t q = 0; q < na; ++q) {
      memcpy(AR + q * n, R + act[q] * n, sizeof(double) * n);
      if (iter > 1) {
        memcpy(APa + q * n, P + act[q] * n, sizeof(double) * n);
        memcpy(AAP + q * n, AP + act[q] * n, sizeof(double) * n);
      }
    }
    apply_M_block(h, n, na, AR, vtmp);
    if (m_y > 0) orthog_y(n, m_y, Y, Uy, na, AR, tmp);
    /* AR <- AR - X (X^T AR) */
    gemm_tn(n, neig, X, na, AR, Csm);
    gemm_nn(n, neig, X, na, Csm, tmp);
    for (int64_t i = 0; i < n * na; ++i) AR[i] -= tmp[i];
    if (b_orthonormalize(n, na, AR, U)) {
      ierr = 2;
      break;
    }
    apply_A_block(A, n, na, AR, AAR);
    if (iter > 1) {
      if (b_orthonormalize(n, na, APa, U)) {
        restart = 1;
      } else {
        right_solve_upper(n, na, AAP, U);
        restart = forced;
      }
    }
    /* Gram matrices, assembled block by block into the m x m arrays. */
    const int32_t m = restart ? neig + na : neig + 2 * na;
    double *blkbuf = Csm;
#define PUT(M, r0, c0, rows, cols, SRC, sym)                                   \
  for (int32_t jj = 0; jj < (cols); ++jj)                                      \
    for (int32_t ii = 0; ii < (rows); ++ii) {                                  \
      double v_ = (SRC)[ii + (int64_t)jj * (rows)];                            \
      (M)[(r0) + ii + (int64_t)((c0) + jj) * m] = v_;                          \
      if (sym) (M)[(c0) + jj + (int64_t)((r0) + ii) * m] = v_;                 \
    }
    gemm_tn(n, neig, X, neig, AX, blkbuf); symmetrize(neig, blkbuf); PUT(gA, 0, 0, neig, neig, blkbuf, 0);
    gemm_tn(n, neig, X, na, AAR, blkbuf); PUT(gA, 0, neig, neig, na, blkbuf, 1);
    gemm_tn(n, na, AR, na, AAR, blkbuf); symmetrize(na, blkbuf); PUT(gA, neig, neig, na, na, blkbuf, 0);
    gemm_tn(n, neig, X, neig, X, blkbuf); symmetrize(neig, blkbuf); PUT(gB, 0, 0, neig, neig, blkbuf, 0);
    gemm_tn(n, neig, X, na, AR, blkbuf); PUT(gB, 0, neig, neig, na, blkbuf, 1);
    gemm_tn(n, na, AR, na, AR, blkbuf); symmetrize(na, blkbuf); PUT(gB, neig, neig, na, na, blkbuf, 0);
    if (!restart) {
      const int32_t c2 = neig + na;
      gemm_tn(n, neig, X, na, AAP, blkbuf); PUT(gA, 0, c2, neig, na, blkbuf, 1);
      gemm_tn(n, na, AR, na, AAP, blkbuf); PUT(gA, neig, c2, na, na, blkbuf, 1);
      gemm_tn(n, na, APa, na, AAP, blkbuf); symmetrize(na, blkbuf); PUT(gA, c2, c2, na, na, blkbuf, 0);
      gemm_tn(n, neig, X, na, APa, blkbuf); PUT(gB, 0, c2, neig, na, blkbuf, 1);
      gemm_tn(n, na, AR, na, APa, blkbuf); PUT(gB, neig, c2, na, na, blkbuf, 1);
      gemm_tn(n, na, APa, na, APa, blkbuf); symmetrize(na, blkbuf); PUT(gB, c2, c2, na, na, blkbuf, 0);
    }
#undef PUT
    if (orc_sygv(m, gA, gB, wall, E)) {
      ierr = 2;
      break;
    }
    /* Select the neig smallest (ascending) or largest (descending). */
    double *Es = (double *)malloc(sizeof(double) * m * neig);
    for (int32_t j = 0; j < neig; ++j) {
      int32_t src = largest ? m - 1 - j : j;
      lam[j] = wall[src];
      HIST(lambda_hist, iter, j) = lam[j];
      memcpy(Es + (int64_t)j * m, E + (int64_t)src * m, sizeof(double) * m);
    }
    /* EX = Es(0:neig), ER = Es(neig:neig+na), EP = Es(neig+na:m) */
    double *EX = (double *)malloc(sizeof(double) * neig * neig);
    double *ER = (double *)malloc(sizeof(double) * na * neig);
    double *EP = (double *)malloc(sizeof(double) * na * neig);
    for (int32_t j = 0; j < neig; ++j) {
      for (int32_t i = 0; i < neig; ++i) EX[i + j * neig] = Es[i + (int64_t)j * m];
      for (int32_t i = 0; i < na; ++i) ER[i + j * na] = Es[neig + i + (int64_t)j * m];
      if (!restart)
        for (int32_t i = 0; i < na; ++i) EP[i + j * na] = Es[neig + na + i + (int64_t)j * m];
    }
    /* pp = AR ER (+ APa EP); app = AAR ER (+ AAP EP) */
    gemm_nn(n, na, AR, neig, ER, P);
    gemm_nn(n, na, AAR, neig, ER, AP);
    if (!restart) {
      gemm_nn(n, na, APa, neig, EP, tmp);
      for (int64_t i = 0; i < n * neig; ++i) P[i] += tmp[i];
      gemm_nn(n, na, AAP, neig, EP, tmp);
      for (int64_t i = 0; i < n * neig; ++i) AP[i] += tmp[i];
    }
    gemm_nn(n, neig, X, neig, EX, tmp);
    for (int64_t i = 0; i < n * neig; ++i) X[i] = P[i] + tmp[i];
    gemm_nn(n, neig, AX, neig, EX, tmp);
    for (int64_t i = 0; i < n * neig; ++i) AX[i] = AP[i] + tmp[i];
    free(Es); free(EX); free(ER); free(EP);
  }
  if (n_active > 0) {
    ++iter;
    for (int32_t j = 0; j < neig; ++j) {
      for (int64_t i = 0; i < n; ++i) R[i + j * n] = AX[i + j * n] - X[i + j * n] * lam[j];
      resn[j] = nrm2(n, R + j * n) / lam[j];
      HIST(resnorm_hist, iter - 1, j) = resn[j];
      HIST(lambda_hist, iter - 1, j) = lam[j];
    }
  }
  {
    double mx = 0.0;
    for (int32_t j = 0; j < neig; ++j) mx = fmax(mx, reslambda_check ? relDelta[j] : resn[j]);
    if (mx > tol) {
      ierr = -1;
      memcpy(X, best, blk);
    }
  }
  /* Final Rayleigh-Ritz without preconditioning (lobpcg.m:420-459). */
  ++iter;
  if (m_y > 0) orthog_y(n, m_y, Y, Uy, neig, X, tmp);
  apply_A_block(A, n, neig, X, AX);
  gemm_tn(n, neig, X, neig, AX, gA);
  symmetrize(neig, gA);
  gemm_tn(n, neig, X, neig, X, gB);
  symmetrize(neig, gB);
  if (orc_sygv(neig, gA, gB, wall, E)) { ierr = 1; goto done; }
  gemm_nn(n, neig, X, neig, E, tmp); memcpy(X, tmp, blk);
  gemm_nn(n, neig, AX, neig, E, tmp); memcpy(AX, tmp, blk);
  for (int32_t j = 0; j < neig; ++j) {
    lam[j] = wall[j];
    for (int64_t i = 0; i < n; ++i) R[i + j * n] = AX[i + j * n] - X[i + j * n] * lam[j];
    resn[j] = nrm2(n, R + j * n) / lam[j];
    HIST(resnorm_hist, iter - 1, j) = resn[j];
    HIST(lambda_hist, iter - 1, j) = lam[j];
  }
done:
  if (lambda_out) memcpy(lambda_out, lam, sizeof(double) * neig);
  if (niter_out) *niter_out = iter - 1;
  if (ierr_out) *ierr_out = ierr;
  free(AX); free(P); free(AP); free(R); free(AR); free(AAR); free(APa); free(AAP);
  free(best); free(tmp); free(vtmp); free(gA); free(gB); free(E); free(wall); free(Csm); free(U);
  free(Uy);
#undef HIST
  return 0;
}

/* ---------------------------------------------------------------------------------- */
/* Adjoint gradient for the finite-volume diffusion operator                           */
/* ---------------------------------------------------------------------------------- */
/* With T_ij = 2 k_i k_j / (k_i + k_j), dT_ij/d ln k_i = T_ij k_j / (k_i + k_j), and the
 * Dirichlet term 2 k_i has derivative 2 k_i, dL/d ln k_i = -lam^T (dA/d ln k_i) x. */
int orc_fv_grad(int32_t N, const double *k, const double *x, const double *lam, double *g) {
  if (N <= 0 || !k || !x || !lam || !g) return -1;
  const int64_t NN = (int64_t)N * N, n = NN * N;
#pragma omp parallel for num_threads(nth()) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const int xi = (int)(i % N), yi = (int)((i / N) % N), zi = (int)(i / NN);
    const double ki = k[i];
    double s = 0.0;
    int64_t nb[6];
    int cnt = 0;
    if (zi > 0) nb[cnt++] = i - NN;
    if (yi > 0) nb[cnt++] = i - N;
    if (xi > 0) nb[cnt++] = i - 1;
    if (xi < N - 1) nb[cnt++] = i + 1;
    if (yi < N - 1) nb[cnt++] = i + N;
    if (zi < N - 1) nb[cnt++] = i + NN;
    for (int q = 0; q < cnt; ++q) {
      const int64_t j = nb[q];
      const double kj = k[j];
      const double t = 2.0 * ki * kj / (ki + kj);
      const double dt = t * kj / (ki + kj);
      s += dt * (lam[i] - lam[j]) * (x[i] - x[j]);
    }
    int nd = (xi == 0) + (xi == N - 1);
    s += nd * 2.0 * ki * lam[i] * x[i];
    g[i] = -s;
  }
  return 0;
}
