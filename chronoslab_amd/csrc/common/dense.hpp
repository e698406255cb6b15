// Small dense FP64 linear algebra on the host for the Rayleigh-Ritz steps of LOBPCG
// (matrices of order <= 48): Cholesky, the symmetric-definite generalized eigenproblem by
// Cholesky reduction + the symmetric QR algorithm (MATLAB's eig(A, B) in
// sources/EigenSolver/lobpcg.m:99, :325 and :437), condition numbers (cond in
// sources/EigenSolver/B_orthonormalize.m:9).  All matrices are column-major.  The loops
// run down contiguous columns with no dependence between iterations, and sums keep a fixed
// order, so the results do not depend on how the compiler vectorizes them.
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace chrdense {

using Mat = std::vector<double>;  // column-major m x m

// In-place lower Cholesky (strict upper part zeroed).  Returns false when not SPD.  Column j
// is updated by the columns k < j in increasing k (axpy form: the inner loops run down
// contiguous columns and carry no dependence).
inline bool chol_lower(int m, double *K) {
  for (int j = 0; j < m; ++j) {
    double *cj = K + (size_t)j * m;
    for (int k = 0; k < j; ++k) {
      const double *ck = K + (size_t)k * m;
      const double ljk = ck[j];
      for (int i = j; i < m; ++i) cj[i] -= ck[i] * ljk;
    }
    if (!(cj[j] > 0.0)) return false;
    const double d = std::sqrt(cj[j]);
    cj[j] = d;
    for (int i = j + 1; i < m; ++i) cj[i] /= d;
    for (int i = 0; i < j; ++i) cj[i] = 0.0;
  }
  return true;
}

inline void symmetrize(int m, double *M) {
  for (int j = 0; j < m; ++j)
    for (int i = j + 1; i < m; ++i) {
      const double s = 0.5 * (M[i + j * m] + M[j + i * m]);
      M[i + j * m] = M[j + i * m] = s;
    }
}

// Householder reduction of a symmetric S (destroyed) to tridiagonal form T, with diagonal
// d and off-diagonal e[i] = T(i+1, i) (e[m-1] = 0); each reflector H_k = I - beta_k v_k v_k^T
// (v_k zero above row k+1) is applied to the trailing block as the symmetric rank-2 update
// S - v w^T - w v^T (Golub & Van Loan, Alg. 8.3.1).  Q (when not null) receives
// H_0 H_1 ... H_{m-3}.
inline void tridiagonalize(int m, double *S, double *d, double *e, double *Q) {
  if (Q)
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) Q[i + j * m] = (i == j) ? 1.0 : 0.0;
  std::vector<double> v(m), p(m), y(m);
  for (int k = 0; k + 2 < m; ++k) {
    // Reflector H = I - beta_k v v^T on indices r0 = k+1..m-1 mapping S[r0.., k] onto e_1.
    const int r0 = k + 1;
    double sig = 0.0;
    for (int i = k + 2; i < m; ++i) sig += S[i + k * m] * S[i + k * m];
    if (sig == 0.0) continue;
    const double x0 = S[r0 + k * m], alpha = std::sqrt(x0 * x0 + sig);
    const double v0 = x0 + (x0 >= 0.0 ? alpha : -alpha);
    v[r0] = v0;
    for (int i = k + 2; i < m; ++i) v[i] = S[i + k * m];
    const double beta_k = 2.0 / (v0 * v0 + sig);
    // p = beta_k S22 v, then w = p - (beta_k/2)(p^T v) v (kept in p); S22 <- S22 - v w^T - w v^T.
    for (int i = r0; i < m; ++i) p[i] = 0.0;
    for (int j = r0; j < m; ++j) {
      const double *cj = S + (size_t)j * m;
      const double vj = v[j];
      for (int i = r0; i < m; ++i) p[i] += cj[i] * vj;
    }
    double pv = 0.0;
    for (int i = r0; i < m; ++i) {
      p[i] *= beta_k;
      pv += p[i] * v[i];
    }
    const double kk = 0.5 * beta_k * pv;
    for (int i = r0; i < m; ++i) p[i] -= kk * v[i];
    for (int j = r0; j < m; ++j) {
      double *cj = S + (size_t)j * m;
      const double vj = v[j], pj = p[j];
      for (int i = r0; i < m; ++i) cj[i] -= v[i] * pj + p[i] * vj;
    }
    const double sub = x0 >= 0.0 ? -alpha : alpha;  // H x = sub e_1
    S[r0 + k * m] = S[k + r0 * m] = sub;
    for (int i = k + 2; i < m; ++i) S[i + k * m] = S[k + i * m] = 0.0;
    if (Q) {  // Q <- Q H = Q - (beta_k Q v) v^T
      for (int i = 0; i < m; ++i) y[i] = 0.0;
      for (int j = r0; j < m; ++j) {
        const double *qj = Q + (size_t)j * m;
        const double vj = v[j];
        for (int i = 0; i < m; ++i) y[i] += qj[i] * vj;
      }
      for (int j = r0; j < m; ++j) {
        double *qj = Q + (size_t)j * m;
        const double bvj = beta_k * v[j];
        for (int i = 0; i < m; ++i) qj[i] -= y[i] * bvj;
      }
    }
  }
  for (int i = 0; i < m; ++i) d[i] = S[i + i * m];
  for (int i = 0; i + 1 < m; ++i) e[i] = 0.5 * (S[i + 1 + i * m] + S[i + (i + 1) * m]);
  if (m > 0) e[m - 1] = 0.0;
}

// Implicit QR steps with the Wilkinson shift and Givens bulge chasing (Alg. 8.3.3) on the
// tridiagonal (d, e), in place: d ends as the eigenvalues (unsorted), e is destroyed; the
// rotations accumulate into Q when it is not null.
inline void tridiag_qr(int m, double *d, double *e, double *Q) {
  const double eps = 2.220446049250313e-16;
  int hi = m - 1, guard = 0;
  while (hi > 0 && guard < 60 * m) {
    if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) --lo;
    if (lo > 0) e[lo - 1] = 0.0;
    ++guard;
    // Wilkinson shift from the trailing 2x2 block of T(lo..hi).
    const double dd = 0.5 * (d[hi - 1] - d[hi]), b = e[hi - 1];
    const double mu = d[hi] - b * b / (dd + (dd >= 0.0 ? 1.0 : -1.0) * std::hypot(dd, b));
    double x = d[lo] - mu, z = e[lo], bulge = 0.0;
    for (int k = lo; k < hi; ++k) {
      // Givens rotation on rows/columns k, k+1 zeroing z against x.  sqrt(x^2 + z^2) where
      // the squares can neither overflow nor lose the larger one to underflow, else hypot.
      const double big = std::max(std::fabs(x), std::fabs(z));
      const double r = big > 1e-150 && big < 1e150 ? std::sqrt(x * x + z * z) : std::hypot(x, z);
      const double ir = r == 0.0 ? 0.0 : 1.0 / r;
      const double c = r == 0.0 ? 1.0 : x * ir, s = z * ir;
      if (k > lo) e[k - 1] = r;
      // Rotate the 2x2 block [d_k e_k; e_k d_k1] and the entries beside it.
      const double dk = d[k], dk1 = d[k + 1], ek = e[k];
      d[k] = c * c * dk + 2.0 * c * s * ek + s * s * dk1;
      d[k + 1] = s * s * dk - 2.0 * c * s * ek + c * c * dk1;
      e[k] = c * s * (dk1 - dk) + (c * c - s * s) * ek;
      if (k + 1 < hi) {
        bulge = s * e[k + 1];
        e[k + 1] *= c;
        x = e[k];
        z = bulge;
      }
      if (Q) {
        double *qk = Q + (size_t)k * m, *qk1 = qk + m;
        for (int i = 0; i < m; ++i) {
          const double a = qk[i], b1 = qk1[i];
          qk[i] = c * a + s * b1;
          qk1[i] = -s * a + c * b1;
        }
      }
    }
  }
}

// Symmetric eigen-decomposition: S (symmetric) is destroyed; w = eigenvalues (unsorted),
// Q = orthonormal eigenvectors (not formed when want_q is false).
inline void sym_eig(int m, double *S, double *w, double *Q, bool want_q = true) {
  std::vector<double> d(m), e(m, 0.0);
  tridiagonalize(m, S, d.data(), e.data(), want_q ? Q : nullptr);
  tridiag_qr(m, d.data(), e.data(), want_q ? Q : nullptr);
  for (int i = 0; i < m; ++i) w[i] = d[i];
}

// C <- L^-1 C L^-T (then symmetrized) for the lower Cholesky factor L of B: the standard
// form of A v = lambda B v.  The forward substitution runs over four columns at a time
// (four independent chains of divisions and updates); every entry is computed by the same
// operations in the same order as one column at a time.
inline void whiten(int m, const double *L, double *C) {
  int j = 0;
  for (; j + 4 <= m; j += 4) {  // C = L^-1 C, four columns at a time
    double *c0 = C + (size_t)j * m, *c1 = c0 + m, *c2 = c1 + m, *c3 = c2 + m;
    for (int q = 0; q < m; ++q) {
      const double *lq = L + (size_t)q * m;
      const double lqq = lq[q];
      c0[q] /= lqq;
      c1[q] /= lqq;
      c2[q] /= lqq;
      c3[q] /= lqq;
      const double a0 = c0[q], a1 = c1[q], a2 = c2[q], a3 = c3[q];
      for (int i = q + 1; i < m; ++i) {
        const double li = lq[i];
        c0[i] -= li * a0;
        c1[i] -= li * a1;
        c2[i] -= li * a2;
        c3[i] -= li * a3;
      }
    }
  }
  for (; j < m; ++j) {
    double *c = C + (size_t)j * m;
    for (int q = 0; q < m; ++q) {
      const double *lq = L + (size_t)q * m;
      c[q] /= lq[q];
      const double cq = c[q];
      for (int i = q + 1; i < m; ++i) c[i] -= lq[i] * cq;
    }
  }
  for (j = 0; j < m; ++j) {  // C = C L^-T: column j from the final columns q < j
    double *cj = C + (size_t)j * m;
    for (int q = 0; q < j; ++q) {
      const double *cq = C + (size_t)q * m;
      const double ljq = L[j + (size_t)q * m];
      for (int i = 0; i < m; ++i) cj[i] -= cq[i] * ljq;
    }
    const double ljj = L[j + (size_t)j * m];
    for (int i = 0; i < m; ++i) cj[i] /= ljj;
  }
  symmetrize(m, C);
}

// A v = lambda B v, A symmetric, B SPD: ascending eigenvalues w, B-orthonormal V.
inline bool sygv(int m, const double *A, const double *B, double *w, double *V) {
  Mat L(B, B + (size_t)m * m), C(A, A + (size_t)m * m), Q((size_t)m * m), wt(m);
  if (!chol_lower(m, L.data())) return false;
  whiten(m, L.data(), C.data());
  sym_eig(m, C.data(), wt.data(), Q.data());
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return wt[a] < wt[b]; });
  // V = L^-T Q[:, ord], all columns together: Wt (row i = entry i of every vector) keeps the
  // back substitution's inner loop on contiguous memory.
  Mat Wt((size_t)m * m);
  for (int jj = 0; jj < m; ++jj) {
    w[jj] = wt[ord[jj]];
    const double *q = Q.data() + (size_t)ord[jj] * m;
    for (int i = 0; i < m; ++i) Wt[(size_t)i * m + jj] = q[i];
  }
  for (int i = m - 1; i >= 0; --i) {
    double *wi = Wt.data() + (size_t)i * m;
    for (int k = i + 1; k < m; ++k) {
      const double lki = L[k + (size_t)i * m];
      const double *wk = Wt.data() + (size_t)k * m;
      for (int jj = 0; jj < m; ++jj) wi[jj] -= lki * wk[jj];
    }
    const double lii = L[i + (size_t)i * m];
    for (int jj = 0; jj < m; ++jj) wi[jj] /= lii;
  }
  for (int jj = 0; jj < m; ++jj)
    for (int i = 0; i < m; ++i) V[i + (size_t)jj * m] = Wt[(size_t)i * m + jj];
  return true;
}

// 2-norm condition number of a symmetric matrix (|eigenvalue| ratio).
inline double cond_sym(int m, const double *M) {
  Mat S(M, M + (size_t)m * m), w(m);
  sym_eig(m, S.data(), w.data(), nullptr, false);
  double mx = 0.0, mn = 1e308;
  for (double x : w) {
    mx = std::max(mx, std::fabs(x));
    mn = std::min(mn, std::fabs(x));
  }
  return mn == 0.0 ? INFINITY : mx / mn;
}

// Inverse of an upper-triangular U (column-major).
inline Mat inv_upper(int m, const double *U) {
  Mat R((size_t)m * m, 0.0);
  for (int j = 0; j < m; ++j) {
    R[j + j * m] = 1.0 / U[j + j * m];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += U[i + k * m] * R[k + j * m];
      R[i + j * m] = -s / U[i + i * m];
    }
  }
  return R;
}

}  // namespace chrdense
